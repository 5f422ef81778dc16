"""CPU tests that pin the oracle (the C restatement under oracle/) independently of itself.

The reference has no golden vectors and its only test cannot run (it calls a one-argument
qr_algorithm! that does not exist), so each oracle function is checked here against what
the mathematics fixes: LAPACK eigenvalues, closed forms, a dense QR step, brute-force
products of dense rotation matrices, and identities a plausible mistake (a swapped c/s,
a wrong sign or index, a dropped bulge) would break.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle


def dense(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


def sep_inputs():
    """The package's seeded input generators (pure numpy; no GPU needed)."""
    import conftest
    return conftest.load_package().inputs


# ---------------------------------------------------------------- givensAlgorithm (DLARTG)

@pytest.mark.parametrize("dtype,tol", [(np.float64, 4e-16), (np.float32, 2e-7)])
def test_givens_identities(dtype, tol):
    r = np.random.default_rng(0)
    for f, g in list(r.standard_normal((200, 2))) + [(3.0, 4.0), (-3.0, 4.0), (4.0, -3.0),
                                                      (-4.0, -3.0), (1e-3, 7.0), (-7.0, 1e-3)]:
        c, s, rr = oracle.givens(f, g, dtype)
        f32, g32 = dtype(f), dtype(g)
        assert abs(c * c + s * s - 1) <= 4 * tol
        assert abs(c * f32 + s * g32 - rr) <= 4 * tol * abs(rr)       # maps [f, g] to [r, 0]
        assert abs(-s * f32 + c * g32) <= 4 * tol * abs(rr)
        if abs(f32) > abs(g32):
            assert c > 0                                             # sign rule
        else:
            assert rr > 0


def test_givens_special_cases():
    assert oracle.givens(2.5, 0.0) == (1.0, 0.0, 2.5)
    assert oracle.givens(-2.5, 0.0) == (1.0, 0.0, -2.5)
    assert oracle.givens(0.0, -3.0) == (0.0, 1.0, -3.0)
    assert oracle.givens(0.0, 0.0) == (1.0, 0.0, 0.0)


@pytest.mark.parametrize("scale", [1e300, 1e-300, 2.0 ** 600, 2.0 ** -600])
def test_givens_extreme_magnitudes(scale):
    f, g = 3.0 * scale, 4.0 * scale
    c, s, r = oracle.givens(f, g)
    assert abs(c - 0.6) < 1e-15 and abs(s - 0.8) < 1e-15
    assert abs(r / (5.0 * scale) - 1) < 1e-15


# ---------------------------------------------------------------- Wilkinson shift

def test_wilkinson_shift_is_eigenvalue_nearest_corner():
    r = np.random.default_rng(1)
    for a, b, c in r.standard_normal((100, 3)) * [10, 3, 10]:
        lam = oracle.wilkinson_shift(a, b, c)
        ev = np.linalg.eigvalsh(np.array([[a, b], [b, c]]))
        near = ev[np.argmin(np.abs(ev - c))]
        assert abs(lam - near) <= 1e-13 * max(1.0, np.abs(ev).max())


# ---------------------------------------------------------------- single_shift!

def test_single_shift_equals_dense_shifted_qr_step():
    """One sweep is the implicit shifted QR step: H' = R Q + mu I where H - mu I = Q R."""
    r = np.random.default_rng(2)
    n = 9
    d, e = r.standard_normal(n), r.standard_normal(n - 1)
    mu = 0.37
    dv, ev = d.copy(), e.copy()
    grid = oracle.identity_grid(n, 2)
    lt = np.zeros(n, dtype=np.int64)
    oracle.single_shift(dv, ev, 1, n, mu, grid, lt)
    Qd, Rd = np.linalg.qr(dense(d, e) - mu * np.eye(n))
    Hd = Rd @ Qd + mu * np.eye(n)
    assert np.allclose(np.diag(Hd), dv, atol=1e-12)
    assert np.allclose(np.abs(np.diag(Hd, 1)), np.abs(ev), atol=1e-12)
    assert np.abs(np.triu(Hd, 2)).max() < 1e-12
    # every index got exactly one rotation, recorded in layer 1
    assert np.array_equal(lt[:-1], np.ones(n - 1)) and lt[-1] == 0
    # the recorded rotations realise the similarity: Q^T T Q = T'
    Q = np.asfortranarray(np.eye(n))
    oracle.bulk_rmul_seq(Q, grid)
    assert np.allclose(Q.T @ dense(d, e) @ Q, dense(dv, ev), atol=1e-12)


def test_single_shift_sub_block_leaves_rest_untouched():
    r = np.random.default_rng(3)
    n = 12
    d, e = r.standard_normal(n), r.standard_normal(n - 1)
    dv, ev = d.copy(), e.copy()
    grid = oracle.identity_grid(n, 3)
    lt = np.zeros(n, dtype=np.int64)
    oracle.single_shift(dv, ev, 4, 9, 0.1, grid, lt)   # 1-based block [4, 9]
    assert np.array_equal(dv[:3], d[:3]) and np.array_equal(dv[9:], d[9:])
    assert np.array_equal(ev[:3], e[:3]) and np.array_equal(ev[9:], e[9:])
    assert np.array_equal(lt, [0, 0, 0, 1, 1, 1, 1, 1, 0, 0, 0, 0])


# ---------------------------------------------------------------- bulk apply

def _dense_rot(n, i, c, s):
    G = np.eye(n)
    G[i, i], G[i, i + 1], G[i + 1, i], G[i + 1, i + 1] = c, -s, s, c
    return G   # A @ G maps columns (i, i+1): a_i' = c a_i + s a_{i+1}


def test_bulk_rmul_seq_equals_dense_product():
    r = np.random.default_rng(4)
    m, n, k = 5, 6, 4
    A = np.asfortranarray(r.standard_normal((m, n)))
    th = r.random((k, n - 1)) * 2 * np.pi
    g = np.stack([np.cos(th), np.sin(th)], -1)
    P = np.eye(n)
    for L in range(k):
        for i in range(n - 1):
            P = P @ _dense_rot(n, i, *g[L, i])
    out = A.copy(order="F")
    oracle.bulk_rmul_seq(out, g)
    assert np.allclose(out, A @ P, atol=1e-14, rtol=0)


def test_bulk_rmul_seq_fma_equals_dense_product_and_plain_form():
    """The fused-multiply-add rotation form: equals the dense product of rotation matrices
    to rounding, differs from the plain form only at rounding level, and leaves identity
    entries (s == 0) exactly untouched."""
    r = np.random.default_rng(31)
    m, n, k = 7, 9, 5
    A = np.asfortranarray(r.standard_normal((m, n)))
    th = r.random((k, n - 1)) * 2 * np.pi
    g = np.ascontiguousarray(np.stack([np.cos(th), np.sin(th)], axis=-1))
    g[2, 3] = (1.0, 0.0)
    D = A.copy()
    for L in range(k):
        for i in range(n - 1):
            c, s = g[L, i]
            Gt = np.eye(n)
            Gt[i, i], Gt[i, i + 1], Gt[i + 1, i], Gt[i + 1, i + 1] = c, -s, s, c
            D = D @ Gt
    F = A.copy(order="F")
    oracle.bulk_rmul_seq_fma(F, g)
    P = A.copy(order="F")
    oracle.bulk_rmul_seq(P, g)
    assert np.abs(F - D).max() <= 1e-14 * np.abs(D).max() * 4
    assert 0 < np.abs(F - P).max() <= 1e-14 * np.abs(D).max() * 4


@pytest.mark.parametrize("n,k", [(4, 2), (8, 6), (10, 8), (34, 32), (100, 98), (2000, 64)])
def test_wave_order_bitwise_equals_layer_major(n, k):
    """src/wave.jl applies the grid in anti-diagonal order; without contraction every
    element sees the same operations, so the result is bit-identical."""
    r = np.random.default_rng(n + k)
    m = 37
    A = np.asfortranarray(r.standard_normal((m, n)))
    th = r.random((k, n - 1)) * 2 * np.pi
    g = np.stack([np.cos(th), np.sin(th)], -1)
    g[r.random((k, n - 1)) < 0.3] = (1.0, 0.0)       # identity slots (skipped blocks)
    a1, a2, a3 = A.copy(order="F"), A.copy(order="F"), A.copy(order="F")
    oracle.bulk_rmul_seq(a1, g)
    oracle.bulk_wave(a2, g)
    oracle.bulk_wave(a3, g, threads=3)
    assert np.array_equal(a1, a2) and np.array_equal(a1, a3)


@pytest.mark.parametrize("n,k", [(9, 4), (10, 3), (10, 10)])
def test_wave_order_preconditions(n, k):
    A = np.asfortranarray(np.ones((3, n)))
    with pytest.raises(AssertionError):
        oracle.bulk_wave(A, oracle.identity_grid(n, k))


# ---------------------------------------------------------------- qr_algorithm!

@pytest.mark.parametrize("n,layers", [(2, 128), (4, 2), (10, 8), (64, 128), (200, 16), (256, 128)])
def test_qr_algorithm_eigenpairs(n, layers):
    r = np.random.default_rng(n)
    d, e = np.arange(1.0, n + 1), r.random(n - 1)
    dv, ev = d.copy(), e.copy()
    Q = np.asfortranarray(np.eye(n))
    conv, it, nrot, nblk = oracle.qr_algorithm(dv, ev, Q, layers)
    assert conv
    ref = sl.eigvalsh_tridiagonal(d, e)
    assert np.abs(np.sort(dv) - ref).max() <= 1e-12 * np.abs(ref).max()
    A = dense(d, e)
    assert np.linalg.norm(A @ Q - Q * dv) <= 1e-13 * np.linalg.norm(A)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 1e-13 * n
    assert np.abs(ev).max() <= 1e-15 * np.abs(d).max() * 4


def test_qr_algorithm_apply_modes():
    """Per-index layer packing (src/bulge_chasing.jl:61-63,137-139) applied layer-major
    equals the reference's wave order bit for bit.  Against the rotations applied one by
    one in creation order it is bitwise equal when no split inside a block "un-deflates"
    (d_i = i spectra); with random spectra a split at index i can disappear after the
    lower block's sweeps changed d_{i+1} (ev is only zeroed at the bottom, :214), and the
    packing then reorders rotations that share a column -- a rounding-level difference
    that the reference itself has (it applies the packed grid)."""
    for d, e, bitwise in (
        (np.arange(1.0, 121.0), np.random.default_rng(8).random(119), True),
        (np.random.default_rng(9).standard_normal(120),
         np.random.default_rng(10).standard_normal(119), False),
    ):
        n = d.shape[0]
        outs = []
        for mode in (0, 1, 2, 3):
            dv, ev = d.copy(), e.copy()
            Q = np.asfortranarray(np.eye(n))
            res = oracle.qr_algorithm(dv, ev, Q, 16, apply_mode=mode)
            outs.append((res, dv, ev, Q))
        (r0, dv0, ev0, Q0), (r1, dv1, ev1, Q1), (r2, dv2, ev2, Q2), (r3, dv3, ev3, Q3) = outs
        assert r0 == r1 == r2 == r3 and np.array_equal(dv0, dv3)
        assert np.abs(Q0 - Q3).max() <= 1e-13
        assert np.array_equal(dv0, dv1) and np.array_equal(dv0, dv2)
        assert np.array_equal(Q0, Q1)
        if bitwise:
            assert np.array_equal(Q0, Q2)
        else:
            assert np.abs(Q0 - Q2).max() <= 1e-13


def test_toeplitz_closed_form():
    n, a, b = 50, 2.0, -1.0
    dv, ev = np.full(n, a), np.full(n - 1, b)
    Q = np.asfortranarray(np.eye(n))
    conv, *_ = oracle.qr_algorithm(dv, ev, Q)
    assert conv
    j = np.arange(1, n + 1)
    lam = a + 2 * b * np.cos(j * np.pi / (n + 1))
    assert np.abs(np.sort(dv) - np.sort(lam)).max() <= 1e-13 * 4
    # eigenvectors sqrt(2/(n+1)) sin(i j pi/(n+1)) up to sign and order
    V = np.sqrt(2 / (n + 1)) * np.sin(np.outer(j, j) * np.pi / (n + 1))
    order = np.argsort(dv)
    Qs = Q[:, order]
    Vs = V[:, np.argsort(lam)]
    assert np.abs(np.abs(np.sum(Qs * Vs, 0)) - 1).max() <= 1e-10


def test_two_by_two_closed_form():
    dv, ev = np.array([2.0, -1.0]), np.array([0.5])
    Q = np.asfortranarray(np.eye(2))
    conv, it, _, _ = oracle.qr_algorithm(dv, ev, Q)
    assert conv and it >= 1
    lam = np.linalg.eigvalsh([[2.0, 0.5], [0.5, -1.0]])
    assert np.allclose(np.sort(dv), lam, atol=1e-15)


def test_diagonal_input_needs_no_sweep():
    dv, ev = np.arange(1.0, 9.0), np.zeros(7)
    Q = np.asfortranarray(np.eye(8))
    conv, it, nrot, nblk = oracle.qr_algorithm(dv, ev, Q)
    assert conv and it == 0 and nrot == 0 and nblk == 0
    assert np.array_equal(Q, np.eye(8))


def test_mpmath_eigenvalues_small():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    n = 16
    r = np.random.default_rng(12)
    d, e = r.standard_normal(n), r.standard_normal(n - 1)
    M = mpmath.matrix(dense(d, e).tolist())
    ref = sorted(float(x) for x in mpmath.eigsy(M)[0])
    dv, ev = d.copy(), e.copy()
    oracle.qr_algorithm(dv, ev, None)
    assert np.abs(np.sort(dv) - ref).max() <= 1e-14 * max(abs(x) for x in ref) * 4


def test_float32_solve():
    n = 64
    r = np.random.default_rng(13)
    d, e = np.arange(1.0, n + 1), r.random(n - 1)
    dv, ev = d.astype(np.float32), e.astype(np.float32)
    Q = np.asfortranarray(np.eye(n, dtype=np.float32))
    conv, *_ = oracle.qr_algorithm(dv, ev, Q)
    assert conv
    ref = sl.eigvalsh_tridiagonal(d, e)
    assert np.abs(np.sort(dv.astype(np.float64)) - ref).max() <= 1e-5 * np.abs(ref).max()


def test_odd_n_asserts_and_maxiter():
    with pytest.raises(AssertionError):
        oracle.qr_algorithm(np.ones(5), np.ones(4), None)
    n = 40
    d, e = np.arange(1.0, n + 1), np.random.default_rng(1).random(n - 1)
    conv, it, _, _ = oracle.qr_algorithm(d.copy(), e.copy(), None, 16, maxiter=5)
    assert not conv and it == 6


@pytest.mark.parametrize("n", [1, 3, 9, 51, 201])
def test_odd_n_extension_pinned(n):
    """allow_odd lifts the reference's iseven(n) assertion (an extension; the reference
    marks it "limitation for now", src/bulge_chasing.jl:155).  Pinned by mathematics:
    LAPACK eigenvalues and residual / orthogonality on random input, and the Toeplitz
    closed form lambda_j = a + 2 b cos(j pi / (n + 1)) with its sine eigenvectors."""
    r = np.random.default_rng(n)
    d, e = np.arange(1.0, n + 1), r.random(n - 1)
    dv, ev = d.copy(), e.copy()
    Q = np.asfortranarray(np.eye(n))
    conv, it, nrot, _ = oracle.qr_algorithm(dv, ev, Q, 16, allow_odd=True)
    assert conv
    if n > 1:
        ref = sl.eigvalsh_tridiagonal(d, e)
        assert np.abs(np.sort(dv) - ref).max() <= 1e-12 * np.abs(ref).max()
        A = dense(d, e)
        assert np.linalg.norm(A @ Q - Q * dv) <= 1e-13 * np.linalg.norm(A)
        assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 1e-13 * n
    else:
        assert it == 0 and nrot == 0 and dv[0] == 1.0 and Q[0, 0] == 1.0
    a, b = 2.0, -1.0
    dv, ev = np.full(n, a), np.full(n - 1, b)
    Q = np.asfortranarray(np.eye(n))
    conv, *_ = oracle.qr_algorithm(dv, ev, Q, allow_odd=True)
    j = np.arange(1, n + 1)
    lam = a + 2 * b * np.cos(j * np.pi / (n + 1))
    assert conv and np.abs(np.sort(dv) - np.sort(lam)).max() <= 1e-13 * 4
    V = np.sqrt(2 / (n + 1)) * np.sin(np.outer(j, j) * np.pi / (n + 1))
    Qs, Vs = Q[:, np.argsort(dv)], V[:, np.argsort(lam)]
    assert np.abs(np.abs(np.sum(Qs * Vs, 0)) - 1).max() <= 1e-10


def test_chase_block_resumes_to_full_solve():
    """Running the outer loop block by block through orc_chase_block reproduces the
    whole-solve eigenvalues and iteration count bit for bit."""
    n = 90
    r = np.random.default_rng(4)
    d, e = np.arange(1.0, n + 1), r.random(n - 1)
    dv1, ev1 = d.copy(), e.copy()
    conv, it, nrot, nblk = oracle.qr_algorithm(dv1, ev1, None, 8)
    dv2, ev2 = d.copy(), e.copy()
    st = np.array([n, n, 0, 0, 0, 0], dtype=np.int64)
    blocks = rots = 0
    while True:
        oracle.chase_block(dv2, ev2, st, 8, np.finfo(float).eps, 100 * n)
        if not st[3]:
            break
        blocks += 1
        rots += st[5]
    assert np.array_equal(dv1, dv2) and st[2] == it and blocks == nblk and rots == nrot


# ---------------------------------------------------------------- backward formula

@pytest.mark.parametrize("n", [2, 3, 12, 40])
def test_backward_dense_formula_matches_finite_differences(n):
    """The dense adjoint formula against central differences (h = 1e-6) of a LAPACK
    solve, on every entry of d and e, for L = w.lam + <G, Q> (d_i = i, e ~ U[0.1, 1))."""
    d, e = sep_inputs().tridiag_linear(n, seed=17, e_low=0.1)
    w, G = sep_inputs().backward_loss_weights(n, seed=17)
    lam, Q = oracle.eig_sorted_signfixed(d, e)
    db, eb = oracle.eig_backward_dense(lam, Q, w, G)
    gd, ge = oracle.fd_gradient(d, e, w, G)
    fd_d = np.array([gd[k] for k in range(n)])
    fd_e = np.array([ge[k] for k in range(n - 1)])
    scale = max(np.abs(fd_d).max(), np.abs(fd_e).max() if n > 1 else 0.0, 1.0)
    assert np.abs(db - fd_d).max() <= 1e-6 * scale
    assert np.abs(eb - fd_e).max() <= 1e-6 * scale


def test_backward_dense_formula_eigenvalue_only_closed_form():
    """With Q_bar = 0 the adjoint of lam_k is q_k q_k^T: d_bar = sum_k w_k q_k^2 elementwise
    and e_bar_j = 2 sum_k w_k q_jk q_(j+1)k (Hellmann-Feynman)."""
    n = 9
    d, e = sep_inputs().tridiag_linear(n, seed=5, e_low=0.1)
    lam, Q = oracle.eig_sorted_signfixed(d, e)
    w = np.linspace(-1.0, 2.0, n)
    db, eb = oracle.eig_backward_dense(lam, Q, w, np.zeros((n, n)))
    assert np.allclose(db, (Q ** 2) @ w, rtol=0, atol=1e-14)
    assert np.allclose(eb, 2 * (Q[:-1] * Q[1:]) @ w, rtol=0, atol=1e-14)


def test_fd_directional_matches_dense_gradient():
    """oracle.fd_directional (the n = 4096 backward check along a direction) equals the
    dense-formula gradient projected on the direction, at a size where both are exact
    to FD accuracy."""
    n = 60
    r = np.random.default_rng(60)
    d, e = np.arange(1.0, n + 1), r.uniform(0.1, 1.0, n - 1)
    w, G = r.standard_normal(n), np.asfortranarray(r.standard_normal((n, n)))
    lam, Q = oracle.eig_sorted_signfixed(d, e)
    gd, ge = oracle.eig_backward_dense(lam, Q, w, G)
    vd, ve = r.standard_normal(n), r.standard_normal(n - 1)
    fd = oracle.fd_directional(d, e, w, G, vd, ve)
    g = gd @ vd + ge @ ve
    assert abs(g - fd) <= 1e-6 * (abs(g) + np.sqrt(gd @ gd + ge @ ge))
    # a planted error in the gradient (one entry off by 1e-3) is caught
    gd2 = gd.copy()
    gd2[7] += 1e-3
    assert abs(gd2 @ vd + ge @ ve - fd) > 1e-6 * (abs(g) + np.sqrt(gd @ gd + ge @ ge))


# ---------------------------------------------------------------- identity skip rule

def _apply_rule(A, g, rule):
    """Layer-major product with one of three identity rules, in plain numpy: 'none' applies
    every rotation, 'each' skips every rotation with s == 0, 'block' is the oracle."""
    A = A.copy(order="F")
    k, nm1, _ = g.shape
    for L in range(k):
        for i in range(nm1):
            c, s = g[L, i]
            if rule == "each" and s == 0:
                continue
            a, b = A[:, i].copy(), A[:, i + 1].copy()
            A[:, i], A[:, i + 1] = a * c + b * s, a * -s + b * c
    return A


def _s_zero_grid(n, k, seed, frac=0.6):
    r = np.random.default_rng(seed)
    th = r.random((k, n - 1)) * 2 * np.pi
    g = np.ascontiguousarray(np.stack([np.cos(th), np.sin(th)], -1))
    mask = r.random((k, n - 1)) < frac
    cval = r.choice([-1.0, 0.5, 1.0, 2.0], size=(k, n - 1))
    g[mask, 0] = cval[mask]
    g[mask, 1] = 0.0
    return g


@pytest.mark.parametrize("n,k", [(4, 2), (10, 4), (12, 10), (34, 32), (100, 98), (2000, 64)])
def test_block_skip_rule_equals_wave_order(n, k):
    """Rotations with s == 0 and c in {-1, 0.5, 1, 2}: the sequential oracle's fused-block
    skip gives the wave restatement's result (src/wave.jl:83,98,116,137,152) bit for bit,
    serial and row-split."""
    g = _s_zero_grid(n, k, seed=n + k)
    A = np.asfortranarray(np.random.default_rng(n).standard_normal((5, n)))
    a1, a2, a3 = A.copy(order="F"), A.copy(order="F"), A.copy(order="F")
    oracle.bulk_rmul_seq(a1, g)
    oracle.bulk_wave(a2, g)
    oracle.bulk_wave(a3, g, threads=3)
    assert np.array_equal(a1, a2) and np.array_equal(a1, a3)


def test_block_skip_rule_is_neither_skip_all_nor_skip_none():
    """The rule is pinned against both plausible misreadings: skipping every s == 0
    rotation, and applying every rotation, each differ from the wave order on a grid with
    (-1, 0) and (0.5, 0) entries; a fully s-zero block holding (-1, 0) changes nothing."""
    n, k = 10, 4
    g = _s_zero_grid(n, k, seed=7)
    A = np.asfortranarray(np.random.default_rng(8).standard_normal((5, n)))
    w = A.copy(order="F")
    oracle.bulk_wave(w, g)
    seq = A.copy(order="F")
    oracle.bulk_rmul_seq(seq, g)
    assert np.array_equal(seq, w)
    assert np.abs(_apply_rule(A, g, "each") - w).max() > 1e-3
    assert np.abs(_apply_rule(A, g, "none") - w).max() > 1e-3
    # block (start_col = 2, layers 1, 2): rows 3, 4 of layer 1 and rows 2, 3 of layer 2
    # (1-based); only it is s-zero, with c = -1 everywhere in it
    g = oracle.identity_grid(n, 2)
    th = np.random.default_rng(9).random((2, n - 1)) * 2 * np.pi
    g[..., 0], g[..., 1] = np.cos(th), np.sin(th)
    for (row, lay) in ((3, 1), (4, 1), (2, 2), (3, 2)):
        g[lay - 1, row - 1] = (-1.0, 0.0)
    ref = A.copy(order="F")
    oracle.bulk_wave(ref, g)
    g2 = g.copy()
    for (row, lay) in ((3, 1), (4, 1), (2, 2), (3, 2)):
        g2[lay - 1, row - 1] = (1.0, 0.0)
    out = A.copy(order="F")
    oracle.bulk_rmul_seq(out, g2)
    assert np.array_equal(ref, out)
