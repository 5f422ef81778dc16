"""GPU parity of the round-based batched solver (no counterpart in the reference).

Every problem of a batch runs qr_algorithm! semantics (src/bulge_chasing.jl:146-288) with
the single solve's device chase, and its block of rotations is applied in wave order as
three shears per rotation (the bulk apply's form).  So each problem's eigenvalues,
off-diagonal, iteration count and convergence flag are compared bit for bit with the
oracle's qr_algorithm (they come from the chase alone), and its eigenvectors to
||Q - Q_oracle||_F <= 1e-13 sqrt(n) (Q has Frobenius norm sqrt(n)); full-size batches are
checked against LAPACK and by residual / orthogonality.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle

pytestmark = pytest.mark.gpu


def _batch(sep, ns, seed):
    ns = np.asarray(ns, dtype=np.int64)
    d, e, off_d, off_q = sep.inputs.batched_problems(ns, seed=seed)
    return ns, d, e, off_d, off_q


def _solve(sep, handle, ns, d, e, off_d, off_q, max_layers=0, tune=False):
    import torch
    tot_q = int(off_q[-1] + ns[-1] ** 2)
    dvd = torch.from_numpy(d.copy()).cuda()
    evd = torch.from_numpy(e.copy()).cuda()
    Q = torch.zeros(tot_q, dtype=torch.float64, device="cuda")
    for p, n in enumerate(ns):
        Q[off_q[p]:off_q[p] + n * n].view(n, n).fill_diagonal_(1.0)
    if tune:
        sep.tune_batched(ns, off_d, off_q, dvd, evd, handle=handle)
    conv, iters = sep.qr_batched_(ns, off_d, off_q, dvd, evd, Q, max_layers, handle=handle)
    return conv, iters, dvd.cpu().numpy(), evd.cpu().numpy(), Q.cpu().numpy()


def _check_oracle(ns, d, e, off_d, off_q, ml, conv, iters, gd, ge, gQ, which=None):
    for p in (range(len(ns)) if which is None else which):
        n, o = int(ns[p]), off_d[p]
        dv, ev = d[o:o + n].copy(), e[o:o + n - 1].copy()
        Qo = np.asfortranarray(np.eye(n))
        oconv, oit, _, _ = oracle.qr_algorithm(dv, ev, Qo, ml, apply_mode=3)
        assert bool(conv[p]) == oconv and iters[p] == oit, (p, n)
        assert np.array_equal(gd[o:o + n], dv), (p, n)
        assert np.array_equal(ge[o:o + n - 1], ev), (p, n)
        Qg = gQ[off_q[p]:off_q[p] + n * n].reshape(n, n).T   # column-major
        assert np.linalg.norm(Qg - Qo) <= 1e-13 * np.sqrt(n), (p, n, np.linalg.norm(Qg - Qo))


@pytest.mark.parametrize("max_layers,fused_depth,rows_per_lane",
                         [(0, 0, 0), (16, 8, 1), (20, 8, 2), (64, 8, 2), (16, 4, 1), (7, 4, 1),
                          (16, 16, 1), (40, 16, 1)])
def test_rounds_bitwise_equal_to_oracle(sep, handle, max_layers, fused_depth, rows_per_lane):
    r = np.random.default_rng(21)
    ns = (2 * r.integers(1, 70, size=40)).astype(np.int64)
    ns[:4] = [2, 4, 130, 200]
    print(f"[batched] max_layers={max_layers} KF={fused_depth} R={rows_per_lane} B={len(ns)} "
          f"nmax={ns.max()}", flush=True)
    ns, d, e, off_d, off_q = _batch(sep, ns, seed=13)
    handle.set_batched_config(fused_depth, rows_per_lane)
    try:
        conv, iters, gd, ge, gQ = _solve(sep, handle, ns, d, e, off_d, off_q, max_layers)
    finally:
        handle.set_batched_config(0, 0)
    ml = max_layers if max_layers > 0 else 16  # the default block length
    _check_oracle(ns, d, e, off_d, off_q, ml, conv, iters, gd, ge, gQ)


def test_rounds_full_size_sample_with_tuner(sep, handle):
    """Batched-solve workload shapes (even orders in [32, 512], d_i = i + N(0, 0.1),
    e ~ U[0,1)): a 512-problem sample; the autotuner runs first.  LAPACK eigenvalues,
    residual and orthogonality for every 8th problem; then a solve with max_layers = 32
    under the tuned apply configuration, bitwise against the oracle for 12 problems."""
    ns = sep.inputs.batched_orders(4096)[:512]
    print(f"[batched] tuner + B={len(ns)}", flush=True)
    ns, d, e, off_d, off_q = _batch(sep, ns, seed=5)
    conv, iters, gd, ge, gQ = _solve(sep, handle, ns, d, e, off_d, off_q, 0, tune=True)
    assert conv.all()
    for p in range(0, len(ns), 8):
        n, o = int(ns[p]), off_d[p]
        lam = gd[o:o + n]
        ref = sl.eigvalsh_tridiagonal(d[o:o + n], e[o:o + n - 1])
        assert np.abs(np.sort(lam) - ref).max() <= 1e-12 * np.abs(ref).max()
        Qg = gQ[off_q[p]:off_q[p] + n * n].reshape(n, n).T
        A = np.diag(d[o:o + n]) + np.diag(e[o:o + n - 1], 1) + np.diag(e[o:o + n - 1], -1)
        assert np.linalg.norm(A @ Qg - Qg * lam) <= 1e-13 * np.linalg.norm(A)
        assert np.linalg.norm(Qg.T @ Qg - np.eye(n)) <= 1e-13 * n
    # the tuned configuration changes only speed: an explicit block length gives the
    # oracle's bits whatever fused depth / rows per lane the tuner chose
    conv, iters, gd, ge, gQ = _solve(sep, handle, ns, d, e, off_d, off_q, 32)
    _check_oracle(ns, d, e, off_d, off_q, 32, conv, iters, gd, ge, gQ, which=range(0, 512, 43))


def test_rounds_repeat_is_deterministic(sep, handle):
    """Two solves of the same batch give identical bits."""
    ns = np.array([12, 64, 100, 8, 256, 30], dtype=np.int64)
    ns, d, e, off_d, off_q = _batch(sep, ns, seed=3)
    a = _solve(sep, handle, ns, d, e, off_d, off_q, 32)
    b = _solve(sep, handle, ns, d, e, off_d, off_q, 32)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_rounds_errors(sep, handle):
    import torch
    z = torch.zeros(64, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        handle.set_batched_config(4, 2)
    with pytest.raises(AssertionError):
        sep.qr_batched_(np.array([4, 5]), np.array([0, 4]), np.array([0, 16]), z, z, z,
                        handle=handle)
    with pytest.raises(ValueError):
        sep.qr_batched_(np.array([4, 0]), np.array([0, 4]), np.array([0, 16]), z, z, z,
                        handle=handle)
    with pytest.raises(ValueError):
        sep.qr_batched_(np.array([4]), np.array([0]), np.array([0]), z, z, z, max_layers=300,
                        handle=handle)
    conv, iters = sep.qr_batched_(np.array([], dtype=np.int64), np.array([], dtype=np.int64),
                                  np.array([], dtype=np.int64), z, z, z, handle=handle)
    assert conv.shape == (0,) and iters.shape == (0,)


def test_graph_replays_match_eager_rounds(sep, monkeypatch):
    """The default path replays a captured graph of four rounds until every problem has
    finished; SEP_BATCHED_GRAPH=0 launches the same rounds one at a time.  Three solves on
    the graph handle reuse the same device buffers (so the second and third replay the
    graph captured by the first), with the inputs rewritten by ordinary torch work between
    them; every solve must equal the eager one bit for bit, and the oracle for a sample."""
    import torch
    r = np.random.default_rng(8)
    ns = (2 * r.integers(1, 100, size=64)).astype(np.int64)
    ns[:3] = [2, 6, 256]
    ns, d, e, off_d, off_q = _batch(sep, ns, seed=17)
    print(f"[batched] graph vs eager, B={len(ns)} nmax={ns.max()}", flush=True)
    monkeypatch.setenv("SEP_BATCHED_GRAPH", "0")
    he = sep.Handle(0)
    monkeypatch.delenv("SEP_BATCHED_GRAPH")
    hg = sep.Handle(0)
    try:
        ref = _solve(sep, he, ns, d, e, off_d, off_q, 16)
        tot_q = int(off_q[-1] + ns[-1] ** 2)
        dvd = torch.empty(len(d), dtype=torch.float64, device="cuda")
        evd = torch.empty(len(e), dtype=torch.float64, device="cuda")
        Q = torch.empty(tot_q, dtype=torch.float64, device="cuda")
        for rep in range(3):
            dvd.copy_(torch.from_numpy(d))
            evd.copy_(torch.from_numpy(e))
            Q.zero_()
            for p, n in enumerate(ns):
                Q[off_q[p]:off_q[p] + n * n].view(n, n).fill_diagonal_(1.0)
            torch.cuda.synchronize()
            conv, iters = sep.qr_batched_(ns, off_d, off_q, dvd, evd, Q, 16, handle=hg)
            out = (conv, iters, dvd.cpu().numpy(), evd.cpu().numpy(), Q.cpu().numpy())
            for x, y in zip(out, ref):
                assert np.array_equal(x, y), rep
        _check_oracle(ns, d, e, off_d, off_q, 16, *ref, which=range(0, len(ns), 9))
    finally:
        he.close()
        hg.close()


def test_eigenvalues_only(sep, handle):
    """Q = None: the same chase without the application.  dv / ev / iteration counts equal
    the eigenvector solve's bit for bit (the chase never reads Q) and LAPACK's eigenvalues."""
    import torch
    r = np.random.default_rng(31)
    ns = (2 * r.integers(1, 128, size=96)).astype(np.int64)
    ns, d, e, off_d, off_q = _batch(sep, ns, seed=23)
    conv, iters, gd, ge, _ = _solve(sep, handle, ns, d, e, off_d, off_q, 32)
    dvd = torch.from_numpy(d.copy()).cuda()
    evd = torch.from_numpy(e.copy()).cuda()
    conv2, iters2 = sep.qr_batched_(ns, off_d, off_q, dvd, evd, None, 32, handle=handle)
    assert np.array_equal(conv, conv2) and np.array_equal(iters, iters2)
    assert np.array_equal(dvd.cpu().numpy(), gd) and np.array_equal(evd.cpu().numpy(), ge)
    for p in range(0, len(ns), 5):
        n, o = int(ns[p]), off_d[p]
        ref = sl.eigvalsh_tridiagonal(d[o:o + n], e[o:o + n - 1])
        assert np.abs(np.sort(gd[o:o + n]) - ref).max() <= 1e-12 * np.abs(ref).max()


def test_odd_orders_extension(sep):
    """Odd orders after sep_set_odd_order (rejected by default, like the reference's
    iseven(n) assertion): bit for bit against the oracle with allow_odd."""
    h = sep.Handle(0)
    try:
        r = np.random.default_rng(41)
        ns = r.integers(2, 160, size=48).astype(np.int64)
        ns[:3] = [3, 5, 257]
        ns, d, e, off_d, off_q = _batch(sep, ns, seed=29)
        with pytest.raises(AssertionError):
            _solve(sep, h, ns, d, e, off_d, off_q, 16)
        h.set_odd_order(True)
        conv, iters, gd, ge, gQ = _solve(sep, h, ns, d, e, off_d, off_q, 16)
        for p in range(len(ns)):
            n, o = int(ns[p]), off_d[p]
            dv, ev = d[o:o + n].copy(), e[o:o + n - 1].copy()
            Qo = np.asfortranarray(np.eye(n))
            oconv, oit, _, _ = oracle.qr_algorithm(dv, ev, Qo, 16, apply_mode=3, allow_odd=True)
            assert bool(conv[p]) == oconv and iters[p] == oit, (p, n)
            assert np.array_equal(gd[o:o + n], dv) and np.array_equal(ge[o:o + n - 1], ev)
            Qg = gQ[off_q[p]:off_q[p] + n * n].reshape(n, n).T
            assert np.linalg.norm(Qg - Qo) <= 1e-13 * np.sqrt(n), (p, n)
    finally:
        h.close()


@pytest.mark.parametrize("ns", [[3072, 1024, 64, 8, 2], [3100, 600, 32]])
def test_rounds_large_orders(sep, handle, ns):
    """Problems up to order 3072 are chased in LDS beside the apply; a batch whose largest
    order exceeds that chases from global memory: both bit for bit against the oracle."""
    ns, d, e, off_d, off_q = _batch(sep, ns, seed=int(sum(ns)))
    conv, iters, gd, ge, gQ = _solve(sep, handle, ns, d, e, off_d, off_q, max_layers=32)
    _check_oracle(ns, d, e, off_d, off_q, 32, conv, iters, gd, ge, gQ)


def _tridiag_residuals(d, e, Q, lam):
    """||T Q - Q diag(lam)||_F for the tridiagonal T(d, e), in O(n^2)."""
    TQ = d[:, None] * Q
    TQ[:-1] += e[:, None] * Q[1:]
    TQ[1:] += e[:, None] * Q[:-1]
    return np.linalg.norm(TQ - Q * lam[None, :])


def test_full_batch_4096_problems(sep, handle):
    """The batched workload at its stated size: 4096 problems of even order in [32, 512]
    (seed 5, as timed), autotuned configuration.  Every problem: converged, eigenvalues
    against LAPACK to 1e-12, ||T Q - Q Lam|| <= 1e-13 ||T||, ||Q^T Q - I|| <= 1e-13 n.
    Then the same batch with the default block length, 32 problems (the largest and 31
    spread over the batch) bit for bit against the oracle."""
    ns = sep.inputs.batched_orders(4096)
    print(f"[batched] full batch B={len(ns)} sum n^2={int((ns * ns).sum())}", flush=True)
    ns, d, e, off_d, off_q = _batch(sep, ns, seed=5)
    conv, iters, gd, ge, gQ = _solve(sep, handle, ns, d, e, off_d, off_q, 0, tune=True)
    assert conv.all()
    worst = [0.0, 0.0, 0.0]
    for p in range(len(ns)):
        n, o = int(ns[p]), off_d[p]
        dd, ee, lam = d[o:o + n], e[o:o + n - 1], gd[o:o + n]
        ref = sl.eigvalsh_tridiagonal(dd, ee)
        Qg = gQ[off_q[p]:off_q[p] + n * n].reshape(n, n).T
        tn = np.sqrt((dd * dd).sum() + 2 * (ee * ee).sum())
        worst[0] = max(worst[0], np.abs(np.sort(lam) - ref).max() / np.abs(ref).max())
        worst[1] = max(worst[1], _tridiag_residuals(dd, ee, Qg, lam) / tn)
        worst[2] = max(worst[2], np.linalg.norm(Qg.T @ Qg - np.eye(n)) / n)
    print(f"[batched] worst eig {worst[0]:.2e} residual {worst[1]:.2e} orth/n {worst[2]:.2e}",
          flush=True)
    assert worst[0] <= 1e-12 and worst[1] <= 1e-13 and worst[2] <= 1e-13
    conv, iters, gd, ge, gQ = _solve(sep, handle, ns, d, e, off_d, off_q, 16)
    which = sorted({int(np.argmax(ns))} | set(range(0, len(ns), 133)))
    assert len(which) >= 32
    _check_oracle(ns, d, e, off_d, off_q, 16, conv, iters, gd, ge, gQ, which=which)
