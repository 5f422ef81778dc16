"""GPU tests of the opt-in fast chase rotation (sep_set_chase_rotations(SEP_CHASE_FAST)).

The fast rotation is not the reference's givensAlgorithm bit for bit (c, s and r within
about an ulp of it, c >= 0 always), so these tests hold it to the accuracy the solve
promises instead of to bitwise equality: eigenvalues against LAPACK to 1e-12 relative,
||AQ - Q Lam|| / ||A|| <= 1e-13, ||Q^T Q - I|| <= 1e-13 n, and eigenvectors equal to
LAPACK's up to column sign where the spectrum is well separated.  Inputs at extreme
magnitudes send chunks of chase steps through the exact rotation (the fast formula's range
test), which must give the same accuracy.
"""
import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fast_handle(sep):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    h = sep.Handle(0)
    h.set_chase_rotations("fast")
    yield h
    h.close()


def _dense(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


def _solve(sep, h, d, e, max_layers=128, Q=True):
    n = d.shape[0]
    H = sep.SymTridiagonal(d.copy(), e.copy())
    Qm = np.asfortranarray(np.eye(n)) if Q else None
    _, conv, it = sep.qr_algorithm_(H, Qm, max_layers, handle=h)
    return H.dv, Qm, conv, it


def _check_solution(d, e, lam, Q, vec_tol=None):
    n = d.shape[0]
    ref, V = sl.eigh_tridiagonal(d, e)
    scale = np.abs(ref).max()
    order = np.argsort(lam, kind="stable")
    assert np.abs(lam[order] - ref).max() <= 1e-12 * scale
    if Q is None:
        return
    A = _dense(d, e)
    assert np.linalg.norm(A @ Q - Q * lam) <= 1e-13 * np.linalg.norm(A)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 1e-13 * n
    if vec_tol is not None:
        Qs = Q[:, order]
        sign = np.sign(np.sum(Qs * V, axis=0))
        assert np.abs(Qs * sign - V).max() <= vec_tol


def test_readme_example_n2000_fast(sep, fast_handle, handle):
    """Configuration 1 with the fast rotation: the same accuracy as the exact chase, and an
    iteration count within a few percent of it."""
    n = 2000
    d, e = sep.inputs.tridiag_linear(n)
    lam, Q, conv, it = _solve(sep, fast_handle, d, e)
    assert conv
    # d_i = i: eigenvalue gaps are O(1), so eigenvectors are determined to ~eps ||A||
    _check_solution(d, e, lam, Q, vec_tol=1e-10)
    _, _, conv_x, it_x = _solve(sep, handle, d, e, Q=False)
    assert conv_x and abs(it - it_x) <= 0.05 * it_x


@pytest.mark.parametrize("n,max_layers,seed", [(2, 128, 1), (10, 8, 2), (64, 128, 3),
                                                 (256, 64, 4), (500, 7, 5), (130, 1, 6)])
def test_fast_matches_lapack(sep, fast_handle, n, max_layers, seed):
    d, e = sep.inputs.tridiag_linear(n, seed=seed)
    lam, Q, conv, _ = _solve(sep, fast_handle, d, e, max_layers)
    assert conv
    _check_solution(d, e, lam, Q)


def test_fast_random_spectrum(sep, fast_handle):
    r = np.random.default_rng(11)
    n = 300
    d, e = r.standard_normal(n), r.standard_normal(n - 1)
    lam, Q, conv, _ = _solve(sep, fast_handle, d, e, 32)
    assert conv
    _check_solution(d, e, lam, Q)


@pytest.mark.parametrize("scale", [2.0 ** -560, 2.0 ** 520])
def test_fast_extreme_magnitudes(sep, fast_handle, scale):
    """f^2 + g^2 leaves (2^-1000, 2^1000): those chunks are rerun with the exact rotation."""
    n = 120
    d, e = sep.inputs.tridiag_linear(n, seed=9)
    d, e = d * scale, e * scale
    lam, Q, conv, _ = _solve(sep, fast_handle, d, e, 64)
    assert conv
    ref = sl.eigvalsh_tridiagonal(d, e)
    assert np.all(np.isfinite(lam))
    assert np.abs(np.sort(lam) - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 1e-13 * n


def test_fast_eigenvalues_only_device(sep, fast_handle):
    import torch
    n = 700
    d, e = sep.inputs.tridiag_linear(n, seed=12)
    H = sep.SymTridiagonal(torch.from_numpy(d.copy()).cuda(), torch.from_numpy(e.copy()).cuda())
    _, conv, _ = sep.qr_algorithm_(H, None, handle=fast_handle)
    assert conv
    _check_solution(d, e, H.dv.cpu().numpy(), None)


def test_mode_switch_back_to_exact_is_bitwise(sep):
    """A handle switched to "fast" and back chases bit for bit like the oracle again."""
    import oracle
    n = 200
    d, e = sep.inputs.tridiag_linear(n, seed=13)
    h = sep.Handle(0)
    try:
        h.set_chase_rotations("fast")
        h.set_chase_rotations("exact")
        od, oe = d.copy(), e.copy()
        oracle.qr_algorithm(od, oe, None, 128)
        H = sep.SymTridiagonal(d.copy(), e.copy())
        sep.qr_algorithm_(H, None, handle=h)
        assert np.array_equal(H.dv, od) and np.array_equal(H.ev, oe)
        with pytest.raises(ValueError):
            h.set_chase_rotations("approximate")
    finally:
        h.close()


def test_float32_ignores_fast_mode(sep, fast_handle, handle):
    """Float32 solves always take the exact rotation: the fast handle's result equals the
    exact handle's bit for bit."""
    n = 128
    d, e = sep.inputs.tridiag_linear(n, seed=14)
    out = []
    for h in (fast_handle, handle):
        H = sep.SymTridiagonal(d.astype(np.float32), e.astype(np.float32))
        Q = np.asfortranarray(np.eye(n, dtype=np.float32))
        _, conv, it = sep.qr_algorithm_(H, Q, 64, handle=h)
        assert conv
        out.append((H.dv, H.ev, Q, it))
    assert out[0][3] == out[1][3]
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a, b)


def _batched(sep, h, ns, d, e, off_d, off_q, max_layers=0):
    import torch
    tot_q = int(off_q[-1] + ns[-1] ** 2)
    dvd = torch.from_numpy(d.copy()).cuda()
    evd = torch.from_numpy(e.copy()).cuda()
    Q = torch.zeros(tot_q, dtype=torch.float64, device="cuda")
    for p, n in enumerate(ns):
        Q[off_q[p]:off_q[p] + n * n].view(n, n).fill_diagonal_(1.0)
    conv, iters = sep.qr_batched_(ns, off_d, off_q, dvd, evd, Q, max_layers, handle=h)
    return conv, iters, dvd.cpu().numpy(), Q.cpu().numpy()


def test_batched_fast_then_exact(sep):
    """The batched solver under the fast rotation: LAPACK eigenvalues, residual and
    orthogonality per problem.  Switched back to "exact", the same handle (whose captured
    round graph is keyed on the mode) gives the oracle's dv and iteration counts bit for
    bit."""
    import oracle
    ns = sep.inputs.batched_orders(4096)[:192].astype(np.int64)
    d, e, off_d, off_q = sep.inputs.batched_problems(ns, seed=5)
    print(f"[fast] batched B={len(ns)} nmax={ns.max()}", flush=True)
    h = sep.Handle(0)
    try:
        h.set_chase_rotations("fast")
        conv, _, gd, gQ = _batched(sep, h, ns, d, e, off_d, off_q)
        assert conv.all()
        for p in range(0, len(ns), 4):
            n, o = int(ns[p]), int(off_d[p])
            lam = gd[o:o + n]
            dd, ee = d[o:o + n], e[o:o + n - 1]
            ref = sl.eigvalsh_tridiagonal(dd, ee)
            assert np.abs(np.sort(lam) - ref).max() <= 1e-12 * np.abs(ref).max(), (p, n)
            Qg = gQ[off_q[p]:off_q[p] + n * n].reshape(n, n).T
            A = _dense(dd, ee)
            assert np.linalg.norm(A @ Qg - Qg * lam) <= 1e-13 * np.linalg.norm(A), (p, n)
            assert np.linalg.norm(Qg.T @ Qg - np.eye(n)) <= 1e-13 * n, (p, n)
        h.set_chase_rotations("exact")
        conv, iters, gd, _ = _batched(sep, h, ns, d, e, off_d, off_q, 32)
        for p in range(0, len(ns), 24):
            n, o = int(ns[p]), int(off_d[p])
            dv, ev = d[o:o + n].copy(), e[o:o + n - 1].copy()
            oconv, oit, _, _ = oracle.qr_algorithm(dv, ev, None, 32)
            assert bool(conv[p]) == oconv and iters[p] == oit, (p, n)
            assert np.array_equal(gd[o:o + n], dv), (p, n)
    finally:
        h.close()
