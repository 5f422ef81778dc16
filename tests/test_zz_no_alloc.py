"""GPU test of the allocation-free apply, sep_bulk_rmul_f64_dev_noalloc (no counterpart in
the reference; for callers that cannot allocate device memory).  It runs the kernels of
sep_bulk_rmul_f64_dev with their scratch in a fixed device array, so through the C ABI its
output must equal that entry point's bit for bit: on two calls in a row, the second with a
smaller input (the array keeps the first call's contents), for every geometry the size rule
picks, and an order beyond the array's capacity is refused without touching A."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(sep, fn, h, A, g):
    import torch
    m, n = A.shape
    rc = getattr(sep.lib(), fn)(h.ptr, m, n, A.data_ptr(), A.stride(1), g.data_ptr(),
                                g.shape[0])
    torch.cuda.synchronize()
    return rc


def _inputs(sep, m, n, k, seed):
    import torch
    A = sep.to_device_colmajor(sep.inputs.gaussian_colmajor(m, n, seed))
    g = torch.from_numpy(sep.inputs.givens_grid(n, k, seed + 1)).cuda()
    return A, g


# (m, n, k): the size rule picks wide (m > 14336), r32 (6144 < m <= 10240), r16deep
# (m <= 6144) and small (k <= 32); each pair is a larger call then a smaller one
PAIRS = [
    ((16384, 2000, 64), (15000, 700, 64)),
    ((8192, 16384, 64), (7000, 3000, 130)),
    ((3000, 9000, 64), (100, 257, 65)),
    ((5000, 4000, 32), (70, 90, 7)),
]


@pytest.mark.parametrize("big,small", PAIRS)
def test_noalloc_equals_allocating_entry(sep, big, small):
    h = sep.Handle(0)
    try:
        for i, (m, n, k) in enumerate((big, small)):
            print(f"[noalloc] call {i}: m={m} n={n} k={k}", flush=True)
            A, g = _inputs(sep, m, n, k, seed=m + n + k)
            B = A.clone()
            assert _call(sep, "sep_bulk_rmul_f64_dev", h, A, g) == sep.SEP_OK
            assert _call(sep, "sep_bulk_rmul_f64_dev_noalloc", h, B, g) == sep.SEP_OK, \
                sep.last_error()
            assert bool((A == B).all().item()), (m, n, k)
    finally:
        h.close()


def test_noalloc_rejects_orders_beyond_its_scratch(sep):
    import torch
    h = sep.Handle(0)
    try:
        m, n, k = 8000, 40000, 64  # the 32-row geometry needs well over 48 MiB here
        A = torch.full((n, m), 3.0, dtype=torch.float64, device="cuda").t()
        g = torch.zeros((k, n - 1, 2), dtype=torch.float64, device="cuda")
        rc = _call(sep, "sep_bulk_rmul_f64_dev_noalloc", h, A, g)
        assert rc == sep.SEP_ERR_ARG
        assert "scratch" in sep.last_error()
        assert bool((A == 3.0).all().item())
        # an ordinary call afterwards still works and matches
        A2, g2 = _inputs(sep, 300, 500, 64, seed=4)
        B2 = A2.clone()
        assert _call(sep, "sep_bulk_rmul_f64_dev", h, A2, g2) == sep.SEP_OK
        assert _call(sep, "sep_bulk_rmul_f64_dev_noalloc", h, B2, g2) == sep.SEP_OK
        assert bool((A2 == B2).all().item())
    finally:
        h.close()
