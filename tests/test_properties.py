"""Property-based parity (hypothesis): random shapes, strides, block lengths and seeds.

CPU part: the oracle's wave-order apply (src/wave.jl:62-162) equals its layer-major product
bit for bit, and the layer-major product equals the dense product of the rotation
matrices, for random (m, n, k) within the wave schedule's preconditions.
GPU part (marked gpu): the HIP apply against the oracle for random (m, n, k, lda, element
type, geometry), and the HIP solve against the oracle bit for bit (dv, ev, iteration and
rotation counts) for random even orders and block lengths, in Float64 and Float32.
Example counts are bounded so that each test runs in seconds; the draws are reproducible
(derandomized).
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

import oracle

CPU_SETTINGS = settings(max_examples=100, deadline=None, derandomize=True,
                        suppress_health_check=[HealthCheck.too_slow])
GPU_SETTINGS = settings(max_examples=150, deadline=None, derandomize=True,
                        suppress_health_check=[HealthCheck.too_slow,
                                               HealthCheck.function_scoped_fixture])


def _grid(k, n, seed, identity_frac=0.2, dtype=np.float64):
    r = np.random.default_rng(seed)
    th = r.random((k, n - 1)) * 2 * np.pi
    g = np.stack([np.cos(th), np.sin(th)], -1)
    g[r.random((k, n - 1)) < identity_frac] = (1.0, 0.0)
    return g.astype(dtype)


@st.composite
def wave_shapes(draw):
    # src/wave.jl:65-66: n - 1 odd (n even), k even, k <= n - 2
    n = 2 * draw(st.integers(2, 60))
    k = 2 * draw(st.integers(1, (n - 2) // 2))
    m = draw(st.integers(1, 70))
    return m, n, k, draw(st.integers(0, 2 ** 31))


@CPU_SETTINGS
@given(wave_shapes())
def test_wave_order_bitwise_equals_layer_major_random(shape):
    m, n, k, seed = shape
    r = np.random.default_rng(seed)
    A = np.asfortranarray(r.standard_normal((m, n)))
    g = _grid(k, n, seed + 1)
    a1, a2 = A.copy(order="F"), A.copy(order="F")
    oracle.bulk_rmul_seq(a1, g)
    oracle.bulk_wave(a2, g)
    assert np.array_equal(a1, a2)


@CPU_SETTINGS
@given(st.integers(2, 24), st.integers(1, 12), st.integers(0, 2 ** 31))
def test_layer_major_equals_dense_product_random(n, k, seed):
    """A * prod_L prod_i G(i,L)^T with G(i,L)^T the n x n matrix mapping columns (i, i+1)
    as a_i' = c a_i + s a_{i+1}, a_{i+1}' = -s a_i + c a_{i+1} (Rotation2 rmul!)."""
    r = np.random.default_rng(seed)
    m = 5
    A = np.asfortranarray(r.standard_normal((m, n)))
    g = _grid(k, n, seed + 1)
    P = np.eye(n)
    for L in range(k):
        for i in range(n - 1):
            c, s = g[L, i]
            Gt = np.eye(n)
            Gt[i, i], Gt[i + 1, i], Gt[i, i + 1], Gt[i + 1, i + 1] = c, s, -s, c
            P = P @ Gt
    ref = A @ P
    out = A.copy(order="F")
    oracle.bulk_rmul_seq(out, g)
    assert np.abs(out - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


# ---------------------------------------------------------------- GPU (HIP library)

GEOMS = [None, "wide", "small", "r16deep", "r16s16", "r16s24"]


@st.composite
def apply_cases(draw):
    n = draw(st.integers(2, 300))
    m = draw(st.integers(1, 300))
    k = draw(st.integers(1, 140))
    pad = draw(st.integers(0, 9))
    dtype = draw(st.sampled_from([np.float64, np.float32]))
    geom = draw(st.sampled_from(GEOMS))
    return m, n, k, pad, dtype, geom, draw(st.integers(0, 2 ** 31))


@pytest.fixture(scope="module")
def geom_handles(sep, handle):
    """One handle per forced apply geometry (SEP_APPLY is read when a handle is created);
    None is the default selection by row count."""
    import os
    made = {None: handle}

    def get(geom):
        if geom not in made:
            old = os.environ.get("SEP_APPLY")
            os.environ["SEP_APPLY"] = geom
            try:
                made[geom] = sep.Handle(0)
            finally:
                if old is None:
                    os.environ.pop("SEP_APPLY", None)
                else:
                    os.environ["SEP_APPLY"] = old
        return made[geom]

    yield get
    for g, h in made.items():
        if g is not None:
            h.close()


@pytest.mark.gpu
@GPU_SETTINGS
@given(apply_cases())
def test_gpu_apply_random_shapes(sep, geom_handles, case):
    """Random m, n, k (any parity, k beyond one pass), a leading dimension padded beyond m
    (padding rows must stay untouched), either element type, every apply geometry: the
    GPU result against the oracle's layer-major product, 1e-14 (f64) / 1e-5 (f32)
    relative."""
    import torch
    m, n, k, pad, dtype, geom, seed = case
    h = geom_handles(geom)
    r = np.random.default_rng(seed)
    A = np.asfortranarray(r.standard_normal((m, n)).astype(dtype))
    g = _grid(k, n, seed + 1, dtype=dtype)
    ref = A.copy(order="F")
    oracle.bulk_rmul_seq(ref, g)
    ld = m + pad
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    big = torch.full((n, ld), 3.0, dtype=tdt, device="cuda")
    big[:, :m] = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    Ad = big.t()[:m, :]
    sep.bulk_wave_order_2x2_rmul_(Ad, torch.from_numpy(g).cuda(), handle=h)
    torch.cuda.synchronize()
    out = np.asfortranarray(Ad.t().contiguous().cpu().numpy().T)
    tol = 1e-14 if dtype == np.float64 else 1e-5
    assert np.linalg.norm(out - ref) <= tol * max(np.linalg.norm(ref), 1e-300), case
    if pad:
        assert torch.all(big[:, m:] == 3.0), case


@pytest.mark.gpu
@GPU_SETTINGS
@given(st.integers(1, 90), st.integers(1, 64), st.integers(0, 2 ** 31))
def test_gpu_solve_random_bitwise(sep, handle, half_n, max_layers, seed):
    """Random even n, block length and seeded input (d_i = i + N(0,1), e ~ U[0,1)): the
    GPU qr_algorithm! equals the oracle's bit for bit on dv, ev and the iteration and
    rotation counts; Q to 1e-13 sqrt(n)."""
    n = 2 * half_n
    r = np.random.default_rng(seed)
    d = np.arange(1.0, n + 1) + r.standard_normal(n)
    e = r.random(n - 1)
    od, oe = d.copy(), e.copy()
    oQ = np.asfortranarray(np.eye(n))
    oconv, oit, onrot, _ = oracle.qr_algorithm(od, oe, oQ, max_layers)
    H = sep.SymTridiagonal(d.copy(), e.copy())
    Q = np.asfortranarray(np.eye(n))
    _, conv, it, nrot = sep.qr_algorithm_(H, Q, max_layers, handle=handle,
                                         return_rotations=True)
    assert conv == oconv and it == oit and nrot == onrot, (n, max_layers, seed)
    assert np.array_equal(H.dv, od) and np.array_equal(H.ev, oe), (n, max_layers, seed)
    assert np.linalg.norm(Q - oQ) <= 1e-13 * np.sqrt(n), (n, max_layers, seed)


@pytest.mark.gpu
@settings(max_examples=60, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.too_slow, HealthCheck.function_scoped_fixture])
@given(st.integers(1, 90), st.integers(1, 64), st.integers(0, 2 ** 31))
def test_gpu_solve_random_bitwise_float32(sep, handle, half_n, max_layers, seed):
    """The same in Float32: the chase's fma, division and square root round once in single
    precision, as the oracle's fmaf, / and sqrtf, so dv, ev and the counts agree bit for
    bit; Q to 1e-5 sqrt(n)."""
    n = 2 * half_n
    r = np.random.default_rng(seed)
    d = (np.arange(1.0, n + 1) + r.standard_normal(n)).astype(np.float32)
    e = r.random(n - 1).astype(np.float32)
    od, oe = d.copy(), e.copy()
    oQ = np.asfortranarray(np.eye(n, dtype=np.float32))
    oconv, oit, onrot, _ = oracle.qr_algorithm(od, oe, oQ, max_layers)
    H = sep.SymTridiagonal(d.copy(), e.copy())
    Q = np.asfortranarray(np.eye(n, dtype=np.float32))
    _, conv, it, nrot = sep.qr_algorithm_(H, Q, max_layers, handle=handle,
                                         return_rotations=True)
    assert conv == oconv and it == oit and nrot == onrot, (n, max_layers, seed)
    assert np.array_equal(H.dv, od) and np.array_equal(H.ev, oe), (n, max_layers, seed)
    assert np.linalg.norm(Q - oQ) <= 1e-5 * np.sqrt(n), (n, max_layers, seed)


@pytest.mark.gpu
@settings(max_examples=30, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.too_slow, HealthCheck.function_scoped_fixture])
@given(st.integers(1, 400), st.integers(0, 7), st.integers(0, 2 ** 31))
def test_gpu_sytrd_orgtr_random(sep, handle, n, pad, seed):
    """Dense -> tridiagonal over random orders (both the unblocked and the blocked factor,
    which switch at n = 256) and padded leading dimensions: d, e against LAPACK dsytrd to
    1e-12 |A|, and Q^T A Q = T, Q^T Q = I with the GPU's own factor."""
    import torch
    from tridiag_ref import lapack_sytrd, sym_matrix, tridiag
    A = sym_matrix(n, seed)
    _, d_ref, e_ref, _ = lapack_sytrd(A)
    ld = n + pad
    big = torch.zeros((n, ld), dtype=torch.float64, device="cuda")
    big[:, :n] = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    Ad = big.t()[:n, :]
    d, e, tau = sep.sytrd_(Ad, handle=handle)
    Q = sep.to_host_colmajor(sep.orgtr(Ad, tau, handle=handle))
    d, e = d.cpu().numpy(), e.cpu().numpy()
    nA = max(np.linalg.norm(A), 1e-300)
    assert np.abs(d - d_ref).max() <= 1e-12 * nA, (n, pad, seed)
    if n > 1:
        assert np.abs(e - e_ref).max() <= 1e-12 * nA, (n, pad, seed)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 4e-14 * n, (n, pad, seed)
    assert np.linalg.norm(Q.T @ A @ Q - tridiag(d, e)) <= 4e-14 * n * nA, (n, pad, seed)
    if pad:
        assert torch.all(big[:, n:] == 0.0), (n, pad, seed)


@pytest.mark.gpu
@settings(max_examples=25, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.too_slow, HealthCheck.function_scoped_fixture])
@given(st.lists(st.integers(1, 80), min_size=1, max_size=24), st.sampled_from([7, 16, 33]),
       st.integers(0, 2 ** 31))
def test_gpu_batched_random_bitwise(sep, handle, halves, max_layers, seed):
    """Random batches (even orders 2..160, 1..24 problems) and block lengths: every
    problem's dv, ev, Q, iteration count and convergence flag equal the oracle's
    qr_algorithm with the two-FMA rotation, bit for bit."""
    import torch
    ns = (2 * np.asarray(halves)).astype(np.int64)
    d, e, off_d, off_q = sep.inputs.batched_problems(ns, seed=seed % 100000)
    tot_q = int(off_q[-1] + ns[-1] ** 2)
    dvd = torch.from_numpy(d.copy()).cuda()
    evd = torch.from_numpy(e.copy()).cuda()
    Q = torch.zeros(tot_q, dtype=torch.float64, device="cuda")
    for p, n in enumerate(ns):
        Q[off_q[p]:off_q[p] + n * n].view(n, n).fill_diagonal_(1.0)
    conv, iters = sep.qr_batched_(ns, off_d, off_q, dvd, evd, Q, max_layers, handle=handle)
    gd, ge, gQ = dvd.cpu().numpy(), evd.cpu().numpy(), Q.cpu().numpy()
    ml = max_layers
    for p, n in enumerate(ns):
        n, o = int(n), off_d[p]
        dv, ev = d[o:o + n].copy(), e[o:o + n - 1].copy()
        Qo = np.asfortranarray(np.eye(n))
        oconv, oit, _, _ = oracle.qr_algorithm(dv, ev, Qo, ml, apply_mode=3)
        assert bool(conv[p]) == oconv and iters[p] == oit, (p, n, ml)
        assert np.array_equal(gd[o:o + n], dv) and np.array_equal(ge[o:o + n - 1], ev), (p, n)
        Qg = gQ[off_q[p]:off_q[p] + n * n].reshape(n, n).T
        assert np.linalg.norm(Qg - Qo) <= 1e-13 * np.sqrt(n), (p, n)


@pytest.mark.gpu
@settings(max_examples=25, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.too_slow, HealthCheck.function_scoped_fixture])
@given(st.integers(2, 300), st.integers(0, 2 ** 31))
def test_gpu_backward_random(sep, handle, n, seed):
    """Random orders (d_i = i + small noise, e in [0.1, 1): well separated spectra) and
    cotangents: the GPU backward against the oracle's dense formula, 1e-11 relative."""
    import torch
    r = np.random.default_rng(seed)
    d = np.arange(1.0, n + 1) + 0.05 * r.standard_normal(n)
    e = 0.1 + 0.9 * r.random(n - 1)
    w = r.standard_normal(n)
    G = np.asfortranarray(r.standard_normal((n, n)))
    lam, Q = oracle.eig_sorted_signfixed(d, e)
    rd, re = oracle.eig_backward_dense(lam, Q, w, G)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    db, eb = sep.eig_backward(dev(lam), sep.to_device_colmajor(Q), dev(w),
                              sep.to_device_colmajor(G), handle=handle)
    db, eb = db.cpu().numpy(), eb.cpu().numpy()
    scale = max(np.abs(rd).max(), np.abs(re).max())
    assert np.abs(db - rd).max() <= 1e-11 * scale, (n, seed)
    assert np.abs(eb - re).max() <= 1e-11 * scale, (n, seed)
