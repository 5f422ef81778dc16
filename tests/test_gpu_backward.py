"""GPU parity of the eigendecomposition backward (no counterpart in the reference).

The FP64-MFMA kernels are compared with the oracle's dense formula
(oracle.eig_backward_dense, itself pinned against central finite differences of a LAPACK
solve in tests/test_oracle.py) and directly against finite differences.  Tolerances: the
GPU and numpy sum the two n^3 products in different orders, so agreement is asked to
1e-11 relative to the largest gradient entry (FP64 with n <= 4096), and to 1e-6 against
finite differences with h = 1e-6.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _dev_colmajor(sep, A):
    return sep.to_device_colmajor(np.asfortranarray(A))


def _check(db, eb, rd, re, tol):
    scale = max(np.abs(rd).max(), np.abs(re).max() if re.size else 0.0, 1e-300)
    assert np.abs(db - rd).max() <= tol * scale
    if re.size:
        assert np.abs(eb - re).max() <= tol * scale


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 17, 64, 127, 128, 129, 200, 300])
def test_backward_matches_dense_formula(sep, handle, n):
    d, e = sep.inputs.tridiag_linear(n, seed=n, e_low=0.1)
    w, G = sep.inputs.backward_loss_weights(n, seed=n)
    lam, Q = oracle.eig_sorted_signfixed(d, e) if n > 1 else (d.copy(), np.ones((1, 1)))
    rd, re = oracle.eig_backward_dense(lam, Q, w, G)
    db, eb = sep.eig_backward(_dev(lam), _dev_colmajor(sep, Q), _dev(w), _dev_colmajor(sep, G),
                              handle=handle)
    _check(db.cpu().numpy(), eb.cpu().numpy(), rd, re, 1e-11)


def test_backward_strided_leading_dimension(sep, handle):
    """Q and Q_bar as column-major views with ld > n."""
    import torch
    n, ld = 150, 173
    d, e = sep.inputs.tridiag_linear(n, seed=3, e_low=0.1)
    w, G = sep.inputs.backward_loss_weights(n, seed=3)
    lam, Q = oracle.eig_sorted_signfixed(d, e)
    rd, re = oracle.eig_backward_dense(lam, Q, w, G)
    Qs = torch.full((n, ld), float("nan"), dtype=torch.float64, device="cuda")
    Gs = torch.full((n, ld), float("nan"), dtype=torch.float64, device="cuda")
    Qv, Gv = Qs.t()[:n, :], Gs.t()[:n, :]   # strides (1, ld)
    Qv.copy_(torch.from_numpy(np.ascontiguousarray(Q)))
    Gv.copy_(torch.from_numpy(np.ascontiguousarray(G)))
    db, eb = sep.eig_backward(_dev(lam), Qv, _dev(w), Gv, handle=handle)
    _check(db.cpu().numpy(), eb.cpu().numpy(), rd, re, 1e-11)


@pytest.mark.parametrize("n", [64, 256])
def test_backward_matches_finite_differences(sep, handle, n):
    """Every entry of d and e: the GPU gradient of L = w.lam + <G, Q> against central
    differences (h = 1e-6) of the sorted, sign-fixed LAPACK eigendecomposition."""
    d, e = sep.inputs.tridiag_linear(n, seed=29, e_low=0.1)
    w, G = sep.inputs.backward_loss_weights(n, seed=29)
    lam, Q = oracle.eig_sorted_signfixed(d, e)
    db, eb = sep.eig_backward(_dev(lam), _dev_colmajor(sep, Q), _dev(w), _dev_colmajor(sep, G),
                              handle=handle)
    gd, ge = oracle.fd_gradient(d, e, w, G)
    fd_d = np.array([gd[k] for k in range(n)])
    fd_e = np.array([ge[k] for k in range(n - 1)])
    _check(db.cpu().numpy(), eb.cpu().numpy(), fd_d, fd_e, 1e-6)


def test_backward_fd_64_random_entries(sep, handle):
    """The differentiated-solve check of the north-star workload: central differences
    (h = 1e-6) on 64 random entries (32 of d, 32 of e) of L = w.lam + <G, Q>, the loss
    evaluated by LAPACK (sorted, sign-fixed), against the GPU backward at the same point.
    n = 1024 (each difference is two full eigendecompositions; the n = 4096 problem is
    checked along a random direction below).  Tolerance 1e-5 of the largest entry
    (measured agreement 8.5e-7)."""
    n = 1024
    d, e = sep.inputs.tridiag_linear(n, seed=1, e_low=0.1)
    w, G = sep.inputs.backward_loss_weights(n, seed=1)
    lam, Q = oracle.eig_sorted_signfixed(d, e)
    db, eb = sep.eig_backward(_dev(lam), _dev_colmajor(sep, Q), _dev(w), _dev_colmajor(sep, G),
                              handle=handle)
    db, eb = db.cpu().numpy(), eb.cpu().numpy()
    rng = np.random.default_rng(1024)
    di = rng.choice(n, 32, replace=False)
    ei = rng.choice(n - 1, 32, replace=False)
    gd, ge = oracle.fd_gradient(d, e, w, G, d_idx=di, e_idx=ei)
    scale = max(np.abs(db).max(), np.abs(eb).max())
    err = max(max(abs(db[k] - gd[k]) for k in di), max(abs(eb[k] - ge[k]) for k in ei))
    assert err <= 1e-5 * scale, (err, scale)


def test_backward_fd_directional_n4096(sep, handle):
    """Full size, n = 4096 (d_i = i, e ~ U[0.1, 1)): grad L . v of the GPU backward against
    the central difference (h = 1e-6) of L along a random unit direction v over all of
    (d, e).  Tolerance 1e-5 |grad L| (measured 4e-7)."""
    n = 4096
    d, e = sep.inputs.tridiag_linear(n, seed=1, e_low=0.1)
    w, G = sep.inputs.backward_loss_weights(n)
    lam, Q = oracle.eig_sorted_signfixed(d, e)
    db, eb = sep.eig_backward(_dev(lam), _dev_colmajor(sep, Q), _dev(w), _dev_colmajor(sep, G),
                              handle=handle)
    db, eb = db.cpu().numpy(), eb.cpu().numpy()
    r = np.random.default_rng(7)
    vd, ve = r.standard_normal(n), r.standard_normal(n - 1)
    nv = np.sqrt(vd @ vd + ve @ ve)
    vd, ve = vd / nv, ve / nv
    fd = oracle.fd_directional(d, e, w, G, vd, ve)
    g = db @ vd + eb @ ve
    gn = np.sqrt(db @ db + eb @ eb)
    assert abs(g - fd) <= 1e-5 * gn, (g, fd, gn)


def test_backward_full_size_from_gpu_solve(sep, handle):
    """The differentiated-solve workload at full size: n = 4096, d_i = i, e ~ U[0.1, 1),
    eigendecomposition by this library's GPU solver (checked by its residual), then the
    backward on that (lam, Q) against the dense formula.  (Finite differences are run at
    n = 64 and 256 above; at n = 4096 each difference would need two full solves.)"""
    import torch
    n = 4096
    d, e = sep.inputs.tridiag_linear(n, seed=1, e_low=0.1)
    w, G = sep.inputs.backward_loss_weights(n)
    H = sep.SymTridiagonal(_dev(d), _dev(e))
    lam_d, Q_d = sep.eigen_decomp_(H, handle=handle)
    torch.cuda.synchronize()
    lam = lam_d.cpu().numpy()
    Q = sep.to_host_colmajor(Q_d)
    A = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.linalg.norm(A @ Q - Q * lam) <= 1e-12 * np.linalg.norm(A)
    db, eb = sep.eig_backward(lam_d, Q_d, _dev(w), _dev_colmajor(sep, G), handle=handle)
    rd, re = oracle.eig_backward_dense(lam, Q, w, G)
    _check(db.cpu().numpy(), eb.cpu().numpy(), rd, re, 1e-11)


def test_backward_batched_matches_single(sep, handle):
    """Packed problems of orders 64..256 (and ragged sizes) against the dense formula."""
    r = np.random.default_rng(9)
    ns = np.concatenate([[1, 2, 7, 128, 129], r.integers(64, 257, size=43)]).astype(np.int64)
    off_d = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    off_q = np.concatenate([[0], np.cumsum(ns * ns)[:-1]]).astype(np.int64)
    lam = np.zeros(int(ns.sum()))
    lamb = np.zeros_like(lam)
    Qp = np.zeros(int((ns * ns).sum()))
    Gp = np.zeros_like(Qp)
    refs = []
    for p, n in enumerate(ns):
        d, e = sep.inputs.tridiag_linear(int(n), seed=100 + p, e_low=0.1)
        w, G = sep.inputs.backward_loss_weights(int(n), seed=100 + p)
        lp, Q = oracle.eig_sorted_signfixed(d, e) if n > 1 else (d.copy(), np.ones((1, 1)))
        o, q = off_d[p], off_q[p]
        lam[o:o + n], lamb[o:o + n] = lp, w
        Qp[q:q + n * n] = Q.ravel(order="F")
        Gp[q:q + n * n] = G.ravel(order="F")
        refs.append(oracle.eig_backward_dense(lp, Q, w, G))
    db, eb = sep.eig_backward_batched(ns, off_d, off_q, _dev(lam), _dev(Qp), _dev(lamb), _dev(Gp),
                                      handle=handle)
    db, eb = db.cpu().numpy(), eb.cpu().numpy()
    for p, n in enumerate(ns):
        o = off_d[p]
        _check(db[o:o + n], eb[o:o + n - 1], refs[p][0], refs[p][1], 1e-11)


def test_backward_errors(sep, handle):
    import torch
    z = torch.zeros(4, dtype=torch.float64, device="cuda")
    Q = torch.zeros(4, 4, dtype=torch.float64, device="cuda")   # row-major: rejected
    with pytest.raises(ValueError):
        sep.eig_backward(z, Q, z, Q, handle=handle)
    with pytest.raises(ValueError):
        sep.eig_backward_batched(np.array([4, 0]), np.array([0, 4]), np.array([0, 16]), z, z,
                                 z, z, handle=handle)


def test_backward_batched_1024_problems(sep, handle):
    """The differentiated-solve batch at its stated size, as timed: 1024 problems of even
    order in [64, 256], solved by the batched solver (eigenvalues in deflation order), N(0,1)
    cotangents; every problem against the dense formula."""
    import torch
    Bb = 1024
    rb = np.random.default_rng(11)
    ns = rb.integers(64, 257, size=Bb).astype(np.int64)
    ns += ns % 2
    d, e, off_d, off_q = sep.inputs.batched_problems(ns, seed=13)
    totq = int(off_q[-1] + ns[-1] ** 2)
    print(f"[backward] batched B={Bb} sum n^3={int((ns.astype(np.float64) ** 3).sum()):.3e}",
          flush=True)
    dv = torch.from_numpy(d.copy()).cuda()
    ev = torch.from_numpy(e.copy()).cuda()
    Q = torch.zeros(totq, dtype=torch.float64, device="cuda")
    for p, n in enumerate(ns):
        Q[off_q[p]:off_q[p] + n * n].view(n, n).fill_diagonal_(1.0)
    conv, _ = sep.qr_batched_(ns, off_d, off_q, dv, ev, Q, handle=handle)
    assert conv.all()
    lbar = rb.standard_normal(dv.shape[0])
    qbar = rb.standard_normal(totq)
    db, eb = sep.eig_backward_batched(ns, off_d, off_q, dv, Q, _dev(lbar), _dev(qbar),
                                      handle=handle)
    lam, Qh, db, eb = dv.cpu().numpy(), Q.cpu().numpy(), db.cpu().numpy(), eb.cpu().numpy()
    for p, n in enumerate(ns):
        o, q = off_d[p], off_q[p]
        Qp = Qh[q:q + n * n].reshape(n, n).T
        Gp = qbar[q:q + n * n].reshape(n, n).T
        rd, re = oracle.eig_backward_dense(lam[o:o + n], Qp, lbar[o:o + n], Gp)
        _check(db[o:o + n], eb[o:o + n - 1], rd, re, 1e-11)
