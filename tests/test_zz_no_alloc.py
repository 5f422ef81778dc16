"""GPU test of the allocation-free apply, sep_bulk_rmul_f64_dev_noalloc (no counterpart in
the reference; for callers that cannot allocate device memory).  It runs the kernels of
sep_bulk_rmul_f64_dev with their scratch in a fixed device array, so through the C ABI its
output must equal that entry point's bit for bit: on two calls in a row, the second with a
smaller input (the array keeps the first call's contents), for every geometry the size rule
picks, and an order beyond the array's capacity is refused without touching A.  Captured
into a graph (its purpose), replays equal the eager calls."""
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


# (m, n, k): on a 256-CU device the size rule picks wide (m > 14336), r16s16 (4096 < m <=
# 8192), r16s24 (m <= 4096) and small (k <= 32); each pair is a larger call then a smaller
# one
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
        m, n, k = 8000, 40000, 64  # the 16-row geometry (r16s16) needs well over 48 MiB here
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


def test_noalloc_under_graph_capture(sep):
    """The entry point's purpose: launched inside a stream capture (global mode, the
    handle's stream set before the capture begins), it records no allocation, memset or
    synchronisation.  One graph holds two applies, a cos/sin grid (scaled shear path) and a
    (rand, rand) grid (five-instruction path, flagged by the column-scale kernel's
    self-resetting completion words); three replays, with eager calls in between, each
    equal the eager results bit for bit."""
    import torch
    h = sep.Handle(0)
    try:
        m, n, k = 3000, 2500, 64
        A0 = sep.to_device_colmajor(sep.inputs.gaussian_colmajor(m, n, 11))
        g1 = torch.from_numpy(sep.inputs.givens_grid(n, k, 12)).cuda()
        g2 = torch.rand((k, n - 1, 2), dtype=torch.float64, device="cuda",
                        generator=torch.Generator(device="cuda").manual_seed(13))
        lib = sep.lib()

        def apply(A, g):
            return lib.sep_bulk_rmul_f64_dev_noalloc(h.ptr, m, n, A.data_ptr(), A.stride(1),
                                                     g.data_ptr(), k)

        h.set_stream(torch.cuda.current_stream().cuda_stream)
        E1, E2 = A0.clone(), A0.clone()
        assert apply(E1, g1) == sep.SEP_OK and apply(E2, g2) == sep.SEP_OK
        torch.cuda.synchronize()
        X1, X2 = A0.clone(), A0.clone()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        torch.cuda.synchronize()
        h.set_stream(s.cuda_stream)
        graph = torch.cuda.CUDAGraph()
        print("[noalloc] capture", flush=True)
        with torch.cuda.graph(graph, stream=s, capture_error_mode="global"):
            rc1 = apply(X1, g1)
            rc2 = apply(X2, g2)
        assert rc1 == sep.SEP_OK and rc2 == sep.SEP_OK, sep.last_error()
        torch.cuda.synchronize()
        for rep in range(3):
            print(f"[noalloc] replay {rep}", flush=True)
            X1.copy_(A0)
            X2.copy_(A0)
            h.set_stream(torch.cuda.current_stream().cuda_stream)
            Y = A0.clone()
            assert apply(Y, g2) == sep.SEP_OK
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(X1, E1) and torch.equal(X2, E2), rep
            assert torch.equal(Y, E2)
        del graph
    finally:
        h.close()
