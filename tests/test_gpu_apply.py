"""GPU parity of the bulk Givens application (bulk_wave_order_2x2_rmul!, src/wave.jl:62).

The HIP apply is compared with the oracle's layer-major sequential Rotation2 rmul!
(src/simple_givens.jl:53-64), which the oracle tests pin bitwise to the reference's wave
order.  Tolerance: ||A_gpu - A_ref||_F / ||A_ref||_F <= 1e-14 (Float64), 1e-5 (Float32);
the two differ only by FMA contraction (the reference's @fastmath kernels contract too).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-14, np.float32: 1e-5}


def _case(m, n, k, dtype, seed, lda=None):
    sep = pytest.importorskip("sep_amd")
    A = sep.inputs.gaussian_colmajor(m, n, seed, dtype)
    g = sep.inputs.givens_grid(n, k, seed + 1, dtype)
    return A, g


def _relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,k", [
    (64, 100, 64), (100, 101, 8), (1, 2, 1), (5, 3, 2), (130, 257, 65), (64, 40, 128),
    (200, 64, 33), (64, 1000, 64), (257, 130, 3), (3, 500, 64),
])
def test_bulk_rmul_host_matches_oracle(sep, handle, dtype, m, n, k):
    A, g = _case(m, n, k, dtype, seed=m * 7 + n * 3 + k)
    ref = A.copy(order="F")
    oracle.bulk_rmul_seq(ref, g)
    out = A.copy(order="F")
    sep.bulk_wave_order_2x2_rmul_(out, g, handle=handle)
    assert _relerr(out, ref) <= TOL[dtype], (m, n, k, _relerr(out, ref))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_bulk_rmul_device_strided(sep, handle, dtype):
    import torch
    m, n, k, ld = 96, 300, 64, 131
    A, g = _case(m, n, k, dtype, seed=5)
    ref = A.copy(order="F")
    oracle.bulk_rmul_seq(ref, g)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    big = torch.full((n, ld), 7.0, dtype=tdt, device="cuda")
    big[:, :m] = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    Ad = big.t()[:m, :]  # strides (1, ld)
    assert Ad.stride() == (1, ld)
    gd = torch.from_numpy(g).cuda()
    sep.bulk_wave_order_2x2_rmul_(Ad, gd, handle=handle)
    torch.cuda.synchronize()
    out = np.asfortranarray(Ad.t().contiguous().cpu().numpy().T)
    assert _relerr(out, ref) <= TOL[dtype]
    # padding rows between m and ld are untouched
    assert torch.all(big[:, m:] == 7.0)


def test_identity_grid_is_exact(sep, handle):
    A, _ = _case(70, 90, 64, np.float64, seed=9)
    out = A.copy(order="F")
    sep.bulk_wave_order_2x2_rmul_(out, oracle.identity_grid(90, 64), handle=handle)
    assert np.array_equal(out, A)


def test_reference_bench_grid(sep, handle):
    """benchmark/benchmark.jl:6 fills every slot with one non-normalised (rand, rand)."""
    r = np.random.default_rng(3)
    m, n, k = 128, 200, 64
    c, s = r.random(), r.random()
    g = np.empty((k, n - 1, 2))
    g[..., 0], g[..., 1] = c, s
    A = np.asfortranarray(r.random((m, n)))
    ref = A.copy(order="F")
    oracle.bulk_rmul_seq(ref, g)
    out = A.copy(order="F")
    sep.bulk_wave_order_2x2_rmul_(out, g, handle=handle)
    assert _relerr(out, ref) <= 1e-13


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_full_size_16384_rows_subset(sep, handle, dtype):
    """m = n = 16384, k = 64: the full benchmark shape on the GPU; rows are independent,
    so 64 sampled rows are checked exactly against the oracle, and every row's 2-norm
    is checked for preservation (rotations are orthogonal)."""
    import torch
    n, k = 16384, 64
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    gen = torch.Generator(device="cuda").manual_seed(11)
    A = torch.randn((n, n), generator=gen, dtype=tdt, device="cuda").t()  # column-major
    g = sep.inputs.givens_grid(n, k, 13, dtype)
    rows = np.sort(np.random.default_rng(1).choice(n, 64, replace=False))
    sub = np.asfortranarray(A[rows, :].cpu().numpy())
    norms0 = torch.linalg.vector_norm(A.double(), dim=1)
    sep.bulk_wave_order_2x2_rmul_(A, torch.from_numpy(g).cuda(), handle=handle)
    torch.cuda.synchronize()
    oracle.bulk_rmul_seq(sub, g)
    got = A[rows, :].cpu().numpy()
    assert _relerr(got, sub) <= TOL[dtype]
    norms1 = torch.linalg.vector_norm(A.double(), dim=1)
    rel = ((norms1 - norms0).abs() / norms0).max().item()
    assert rel <= (1e-13 if dtype == np.float64 else 1e-4)


GEOMETRIES = ["wide", "small", "r16deep", "r16s16", "r16s24"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("geom", GEOMETRIES)
def test_every_geometry_matches_oracle(sep, monkeypatch, geom, dtype):
    """Each apply geometry (forced through SEP_APPLY; the default picks one by the number
    of rows) against the oracle: ragged row counts (not a multiple of the block), n smaller
    than the pipeline depth, k beyond one pass, a non-rotation grid (general path), and an
    identity grid, which must leave A bit for bit unchanged."""
    monkeypatch.setenv("SEP_APPLY", geom)
    tol = TOL[dtype]
    h = sep.Handle(0)
    try:
        for m, n, k in ((37, 9, 64), (70, 150, 64), (33, 70, 130), (2000, 300, 64), (65, 2, 3)):
            A, g = _case(m, n, k, dtype, seed=m + n + k)
            ref = A.copy(order="F")
            oracle.bulk_rmul_seq(ref, g)
            out = A.copy(order="F")
            sep.bulk_wave_order_2x2_rmul_(out, g, handle=h)
            assert _relerr(out, ref) <= tol, (geom, m, n, k, _relerr(out, ref))
        r = np.random.default_rng(4)
        A = np.asfortranarray(r.random((41, 50)).astype(dtype))
        g = r.random((64, 49, 2)).astype(dtype)
        ref = A.copy(order="F")
        oracle.bulk_rmul_seq(ref, g)
        out = A.copy(order="F")
        sep.bulk_wave_order_2x2_rmul_(out, g, handle=h)
        assert _relerr(out, ref) <= 10 * tol, (geom, "general", _relerr(out, ref))
        A, _ = _case(45, 90, 64, dtype, seed=9)
        out = A.copy(order="F")
        sep.bulk_wave_order_2x2_rmul_(out, oracle.identity_grid(90, 64, dtype), handle=h)
        assert np.array_equal(out, A), geom
    finally:
        h.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_config2_n2000_whole_matrix(sep, handle, dtype):
    """The reference's published shape, bench(T, 2000, 2000, 64): the whole matrix against
    the oracle's layer-major product."""
    m = n = 2000
    k = 64
    A, g = _case(m, n, k, dtype, seed=2000)
    ref = A.copy(order="F")
    oracle.bulk_rmul_seq(ref, g)
    out = A.copy(order="F")
    sep.bulk_wave_order_2x2_rmul_(out, g, handle=handle)
    assert _relerr(out, ref) <= TOL[dtype], _relerr(out, ref)


def _s_zero_grid(n, k, dtype, seed, cvals):
    """Proper rotations with 60% of the slots replaced by (c, 0), c drawn from cvals."""
    r = np.random.default_rng(seed)
    th = r.random((k, n - 1)) * 2 * np.pi
    g = np.stack([np.cos(th), np.sin(th)], -1)
    mask = r.random((k, n - 1)) < 0.6
    c = r.choice(cvals, size=(k, n - 1))
    g[mask, 0] = c[mask]
    g[mask, 1] = 0.0
    return np.ascontiguousarray(g.astype(dtype))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("geom", GEOMETRIES)
def test_s_zero_entries_follow_the_reference_block_skip(sep, monkeypatch, geom, dtype):
    """Entries (c, 0) with c != 1: the reference applies them (scaling or negating their
    columns) unless their whole fused block is s-zero (src/wave.jl:83,98,116,137,152).
    Grids with c in {-1, 1} are rotations (three-shear path, the sign goes into the column
    sign tracking); with 0.5 and 2 they are not (five-instruction path).  Checked against
    the oracle, whose rule is pinned bit for bit to the wave restatement."""
    monkeypatch.setenv("SEP_APPLY", geom)
    h = sep.Handle(0)
    try:
        for cvals in ([-1.0, 1.0], [-1.0, 0.5, 1.0, 2.0]):
            for m, n, k in ((70, 150, 64), (33, 71, 130), (64, 10, 8)):
                A = sep.inputs.gaussian_colmajor(m, n, m + n, dtype)
                g = _s_zero_grid(n, k, dtype, seed=n + k, cvals=cvals)
                ref = A.copy(order="F")
                oracle.bulk_rmul_seq(ref, g)
                out = A.copy(order="F")
                sep.bulk_wave_order_2x2_rmul_(out, g, handle=h)
                assert _relerr(out, ref) <= TOL[dtype], (geom, cvals, m, n, k,
                                                         _relerr(out, ref))
    finally:
        h.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("delta_eps", [1.5, -1.5, 40.0, -40.0, 1e4, -1e4])
def test_near_normalised_grids(sep, handle, dtype, delta_eps):
    """Pairs with c^2 + s^2 = 1 + delta, delta of one sign for every rotation: just inside
    and well beyond the residual of correctly rounded cos/sin pairs (0.71 eps), up to 1e4
    eps.  A determinant-one (plain shear) realisation would drift by about k delta; the
    kernel folds each pair's sign(c) sqrt(c^2 + s^2) into per-column scales, so every such
    grid must match the oracle to the usual tolerance."""
    m, n, k = 64, 400, 64
    eps = np.finfo(dtype).eps
    r = np.random.default_rng(17)
    th = r.random((k, n - 1)) * 2 * np.pi
    # pick the scale in float64 so that the rounded pairs keep a residual of one sign
    f = np.sqrt(1.0 + 2 * delta_eps * eps)
    g = np.ascontiguousarray(np.stack([np.cos(th) * f, np.sin(th) * f], -1).astype(dtype))
    # residual in extended precision (its products are exact to 2^-64)
    c, s = g[..., 0].astype(np.longdouble), g[..., 1].astype(np.longdouble)
    res = c * c + s * s - 1
    assert np.all(np.sign(res) == np.sign(delta_eps)) and np.abs(res).max() < 1 / 64
    A = sep.inputs.gaussian_colmajor(m, n, 3, dtype)
    ref = A.copy(order="F")
    oracle.bulk_rmul_seq(ref, g)
    out = A.copy(order="F")
    sep.bulk_wave_order_2x2_rmul_(out, g, handle=handle)
    assert _relerr(out, ref) <= TOL[dtype], _relerr(out, ref)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_grid_beyond_the_band_takes_the_exact_path(sep, handle, dtype):
    """c^2 + s^2 = 1.05 for every pair (beyond the 1/64 band of the scaled shear form):
    the five-instruction path applies the pairs as given."""
    m, n, k = 64, 400, 64
    r = np.random.default_rng(19)
    th = r.random((k, n - 1)) * 2 * np.pi
    f = np.sqrt(1.05)
    g = np.ascontiguousarray(np.stack([np.cos(th) * f, np.sin(th) * f], -1).astype(dtype))
    A = sep.inputs.gaussian_colmajor(m, n, 5, dtype)
    ref = A.copy(order="F")
    oracle.bulk_rmul_seq(ref, g)
    out = A.copy(order="F")
    sep.bulk_wave_order_2x2_rmul_(out, g, handle=handle)
    assert _relerr(out, ref) <= 10 * TOL[dtype], _relerr(out, ref)
