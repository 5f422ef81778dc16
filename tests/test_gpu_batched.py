"""GPU parity of the batched persistent solver (no counterpart in the reference).

Every problem of a batch runs qr_algorithm! semantics (src/bulge_chasing.jl:146-288) with
the same device chase as the single solve, so each problem's eigenvalues and iteration
count are compared bit for bit with the oracle's qr_algorithm on that problem, and its
eigenvectors against the oracle's Q (tolerance: the in-wave apply uses the three-shear
form) and through the residual / orthogonality bounds.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="written for the persistent batched kernel, which was removed "
                     "after a GPU run of test_batched_matches_oracle_per_problem[0] did not "
                     "complete with no cause found (the same device code passed under host "
                     "emulation); the round-based batched solver that replaced it is tested in "
                     "test_gpu_batched_rounds.py"),
]


def _pack(sep, ns, seed):
    d, e, off_d, off_q = sep.inputs.batched_problems(ns, seed=seed)
    return d, e, off_d, off_q


def _run(sep, handle, ns, d, e, off_d, off_q, max_layers=0, tune=False):
    import torch
    tot_q = int(off_q[-1] + ns[-1] ** 2)
    dvd = torch.from_numpy(d.copy()).cuda()
    evd = torch.from_numpy(e.copy()).cuda()
    Q = torch.zeros(tot_q, dtype=torch.float64, device="cuda")
    for p, n in enumerate(ns):
        Q[off_q[p]:off_q[p] + n * n].view(n, n).copy_(torch.eye(n, dtype=torch.float64))
    if tune:
        sep.tune_batched(ns, off_d, off_q, dvd, evd, handle=handle)
    conv, iters = sep.qr_batched_(ns, off_d, off_q, dvd, evd, Q, max_layers, handle=handle)
    return conv, iters, dvd.cpu().numpy(), Q.cpu().numpy()


@pytest.mark.parametrize("max_layers", [0, 16, 64])
def test_batched_matches_oracle_per_problem(sep, handle, max_layers):
    r = np.random.default_rng(3)
    ns = (2 * r.integers(1, 40, size=48)).astype(np.int64)   # even orders 2..78
    ns[:3] = [2, 4, 6]
    d, e, off_d, off_q = _pack(sep, ns, seed=11)
    conv, iters, gd, gQ = _run(sep, handle, ns, d, e, off_d, off_q, max_layers)
    ml = max_layers if max_layers > 0 else 32
    for p, n in enumerate(ns):
        o = off_d[p]
        dv, ev = d[o:o + n].copy(), e[o:o + n - 1].copy()
        Qo = np.asfortranarray(np.eye(n))
        oconv, oit, _, _ = oracle.qr_algorithm(dv, ev, Qo, ml)
        assert conv[p] == oconv and iters[p] == oit, (p, n)
        assert np.array_equal(gd[o:o + n], dv), (p, n)
        Qg = gQ[off_q[p]:off_q[p] + n * n].reshape(n, n).T   # column-major
        assert np.linalg.norm(Qg - Qo) <= 1e-13 * n, (p, n)


def test_batched_config4_sample_full_sizes(sep, handle):
    """Configuration 4 shapes: even orders in [32, 512], d_i = i + N(0, 0.1), e ~ U[0,1).
    A 256-problem sample of the 4096-problem batch is solved and checked against LAPACK
    eigenvalues and by residual / orthogonality; the autotuner is exercised first."""
    ns = sep.inputs.batched_orders(4096)[:256]
    d, e, off_d, off_q = _pack(sep, ns, seed=5)
    conv, iters, gd, gQ = _run(sep, handle, ns, d, e, off_d, off_q, 0, tune=True)
    assert conv.all()
    for p in range(0, len(ns), 17):
        n, o = ns[p], off_d[p]
        lam = gd[o:o + n]
        ref = sl.eigvalsh_tridiagonal(d[o:o + n], e[o:o + n - 1])
        assert np.abs(np.sort(lam) - ref).max() <= 1e-12 * np.abs(ref).max()
        Qg = gQ[off_q[p]:off_q[p] + n * n].reshape(n, n).T
        A = np.diag(d[o:o + n]) + np.diag(e[o:o + n - 1], 1) + np.diag(e[o:o + n - 1], -1)
        assert np.linalg.norm(A @ Qg - Qg * lam) <= 1e-13 * np.linalg.norm(A)
        assert np.linalg.norm(Qg.T @ Qg - np.eye(n)) <= 1e-13 * n


def test_batched_graph_replay_is_stable(sep, handle):
    """The captured graph is replayed for an identical layout and identical buffers;
    every replay, with eager work and synchronisation in between, must give the result
    of the first (captured-and-launched) call."""
    import torch
    ns = np.array([10, 32, 64, 8, 100], dtype=np.int64)
    d, e, off_d, off_q = _pack(sep, ns, seed=2)
    tot_q = int(off_q[-1] + ns[-1] ** 2)
    dvd = torch.empty(d.shape[0], dtype=torch.float64, device="cuda")
    evd = torch.empty(e.shape[0], dtype=torch.float64, device="cuda")
    Q = torch.empty(tot_q, dtype=torch.float64, device="cuda")
    ref = None
    for rep in range(4):
        dvd.copy_(torch.from_numpy(d))
        evd.copy_(torch.from_numpy(e))
        Q.zero_()
        for p, n in enumerate(ns):
            Q[off_q[p]:off_q[p] + n * n].view(n, n).fill_diagonal_(1.0)
        torch.cuda.synchronize()
        conv, iters = sep.qr_batched_(ns, off_d, off_q, dvd, evd, Q, 32, handle=handle)
        out = (conv, iters, dvd.cpu().numpy(), Q.cpu().numpy())
        if ref is None:
            ref = out
        else:
            for a, b in zip(out, ref):
                assert np.array_equal(a, b), rep


def test_batched_errors(sep, handle):
    import torch
    z = torch.zeros(16, dtype=torch.float64, device="cuda")
    with pytest.raises(AssertionError):
        sep.qr_batched_(np.array([4, 5]), np.array([0, 4]), np.array([0, 16]), z, z, z,
                        handle=handle)
