"""CPU oracle for the eigensolver hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this package, and only as the checker (or as the timed CPU baseline).  The HIP
library and its Python host layer never import it.

The C restatement lives in ``sep_oracle.c`` / ``sep_oracle_impl.h`` and follows the
reference files named there line by line.  This module wraps it with ctypes and numpy.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict[str, C.CDLL] = {}


def build(force: bool = False) -> None:
    """Compile both oracle variants with the committed Makefile."""
    args = ["make", "-C", _HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def _lib(fast: bool = False) -> C.CDLL:
    name = "libsep_oracle_fast.so" if fast else "libsep_oracle.so"
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        _declare(lib, "fast_" if fast else "")
        _LIBS[name] = lib
    return _LIBS[name]


_P = C.c_void_p
_I64 = C.c_int64


def _declare(lib: C.CDLL, pfx: str) -> None:
    for suf, ct in (("_f64", C.c_double), ("_f32", C.c_float)):
        f = getattr(lib, pfx + "orc_givens" + suf)
        f.argtypes = [ct, ct, _P, _P, _P]
        f.restype = None
        f = getattr(lib, pfx + "orc_wilkinson_shift" + suf)
        f.argtypes = [ct, ct, ct]
        f.restype = ct
        f = getattr(lib, pfx + "orc_single_shift" + suf)
        f.argtypes = [_I64, _P, _P, _I64, _I64, ct, _P, _P, _P]
        f.restype = None
        f = getattr(lib, pfx + "orc_bulk_rmul_seq" + suf)
        f.argtypes = [_I64, _I64, _P, _I64, _P, _I64]
        f.restype = None
        f = getattr(lib, pfx + "orc_bulk_rmul_seq_fma" + suf)
        f.argtypes = [_I64, _I64, _P, _I64, _P, _I64]
        f.restype = None
        f = getattr(lib, pfx + "orc_apply_log" + suf)
        f.argtypes = [_I64, _P, _I64, _P, _P, _P, _I64]
        f.restype = None
        f = getattr(lib, pfx + "orc_bulk_wave" + suf)
        f.argtypes = [_I64, _I64, _P, _I64, _P, _I64]
        f.restype = C.c_int
        f = getattr(lib, pfx + "orc_bulk_wave_parallel" + suf)
        f.argtypes = [_I64, _I64, _P, _I64, _P, _I64, C.c_int]
        f.restype = C.c_int
        f = getattr(lib, pfx + "orc_qr_algorithm" + suf)
        f.argtypes = [_I64, _P, _P, _P, _I64, _I64, ct, _I64, _P, _P, _P, C.c_int, _P, C.c_int]
        f.restype = C.c_int
        f = getattr(lib, pfx + "orc_chase_block" + suf)
        f.argtypes = [_I64, _P, _P, _I64, ct, _I64, _P, _P, _P]
        f.restype = None


def _suf(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "_f64"
    if dt == np.float32:
        return "_f32"
    raise TypeError(f"oracle supports float64/float32, got {dt}")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _fn(name: str, dtype, fast: bool = False):
    return getattr(_lib(fast), ("fast_" if fast else "") + name + _suf(dtype))


def givens(f, g, dtype=np.float64):
    """(c, s, r) of LinearAlgebra.givensAlgorithm(f, g)."""
    ct = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
    c, s, r = ct(), ct(), ct()
    _fn("orc_givens", dtype)(f, g, C.byref(c), C.byref(s), C.byref(r))
    return c.value, s.value, r.value


def wilkinson_shift(h11, h12, h22, dtype=np.float64) -> float:
    return _fn("orc_wilkinson_shift", dtype)(h11, h12, h22)


def identity_grid(n: int, k: int, dtype=np.float64) -> np.ndarray:
    """(n-1) x k grid of identity rotations, stored as [k, n-1, 2] (column-major
    Matrix{Tuple{T,T}} memory order: c, s interleaved, index fastest)."""
    g = np.zeros((k, n - 1, 2), dtype=dtype)
    g[:, :, 0] = 1
    return g


def single_shift(dv, ev, frm: int, to: int, mu, grid, lt, log=False):
    """single_shift! on 1-based block [frm, to]; mutates dv, ev, grid, lt."""
    n = dv.shape[0]
    assert grid.flags.c_contiguous and lt.dtype == np.int64
    _fn("orc_single_shift", dv.dtype)(n, _ptr(dv), _ptr(ev), frm, to, mu, _ptr(grid), _ptr(lt), None)


def bulk_rmul_seq(A: np.ndarray, grid: np.ndarray) -> None:
    """Layer-major sequential rmul! of grid [k, n-1, 2] onto column-major A (m x n)."""
    assert A.flags.f_contiguous and grid.flags.c_contiguous and A.dtype == grid.dtype
    m, n = A.shape
    k = grid.shape[0]
    assert grid.shape[1] == n - 1
    _fn("orc_bulk_rmul_seq", A.dtype)(m, n, _ptr(A), m, _ptr(grid), k)


def bulk_rmul_seq_fma(A: np.ndarray, grid: np.ndarray) -> None:
    """Layer-major product with a1' = fma(a1, c, a2 s), a2' = fma(a1, -s, a2 c)."""
    assert A.flags.f_contiguous and grid.flags.c_contiguous and A.dtype == grid.dtype
    m, n = A.shape
    k = grid.shape[0] if grid.size else 0
    _fn("orc_bulk_rmul_seq_fma", A.dtype)(m, n, _ptr(A), m, _ptr(grid), k)


def apply_log(A: np.ndarray, idx, c, s) -> None:
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    c = np.ascontiguousarray(c, dtype=A.dtype)
    s = np.ascontiguousarray(s, dtype=A.dtype)
    m, _ = A.shape
    _fn("orc_apply_log", A.dtype)(m, _ptr(A), m, _ptr(idx), _ptr(c), _ptr(s), idx.shape[0])


def bulk_wave(A: np.ndarray, grid: np.ndarray, threads: int = 1, fast: bool = False) -> None:
    """bulk_wave_order_2x2_rmul[_parallel]! (requires n-1 odd, k even, k <= n-2)."""
    assert A.flags.f_contiguous and grid.flags.c_contiguous and A.dtype == grid.dtype
    m, n = A.shape
    k = grid.shape[0]
    if threads <= 1:
        rc = _fn("orc_bulk_wave", A.dtype, fast)(m, n, _ptr(A), m, _ptr(grid), k)
    else:
        rc = _fn("orc_bulk_wave_parallel", A.dtype, fast)(m, n, _ptr(A), m, _ptr(grid), k, threads)
    if rc != 0:
        raise AssertionError("wave schedule precondition violated (n-1 odd, k even, k <= n-2)")


def qr_algorithm(dv, ev, Q=None, max_layers: int = 128, tol=None, maxiter=None, apply_mode: int = 0,
                 fast: bool = False, allow_odd: bool = False):
    """qr_algorithm!(H, Q, max_layers, tol, maxiter); mutates dv, ev, Q.

    Returns (converged, iter, n_rotations, n_blocks).  apply_mode 0 = layer-major grid,
    1 = wave order, 2 = chronological rotation log, 3 = layer-major with the rotation as
    two fused multiply-adds.  fast=True runs the FMA-contracted build
    (the timed CPU baseline; parity checks use the default build).  allow_odd accepts odd
    n (the reference asserts iseven(n); the chase itself has no parity condition)."""
    dt = dv.dtype
    n = dv.shape[0]
    if tol is None:
        tol = float(np.finfo(dt).eps)
    if maxiter is None:
        maxiter = 100 * n
    conv = C.c_int32()
    it = C.c_int64()
    nrot = C.c_int64()
    nblk = C.c_int64()
    qp = None
    ldq = n
    if Q is not None:
        assert Q.flags.f_contiguous and Q.dtype == dt and Q.shape == (n, n)
        qp = _ptr(Q)
    rc = _fn("orc_qr_algorithm", dt, fast)(n, _ptr(dv), _ptr(ev), qp, ldq, max_layers, tol, maxiter,
                                     C.byref(conv), C.byref(it), C.byref(nrot), apply_mode,
                                     C.byref(nblk), int(bool(allow_odd)))
    if rc == -1:
        raise AssertionError("qr_algorithm!: n must be even")
    if rc != 0:
        raise AssertionError(f"oracle qr_algorithm failed rc={rc}")
    return bool(conv.value), it.value, nrot.value, nblk.value


def chase_block(dv, ev, state, max_layers: int, tol, maxiter: int):
    """One outer iteration of qr_algorithm! without the apply.  state is int64[6]
    (to, last_block_ends_at, iter, did, early, n_rot); returns the filled grid."""
    n = dv.shape[0]
    grid = identity_grid(n, max_layers, dv.dtype)
    lt = np.zeros(n, dtype=np.int64)
    _fn("orc_chase_block", dv.dtype)(n, _ptr(dv), _ptr(ev), max_layers, tol, maxiter,
                                     _ptr(grid), _ptr(lt), _ptr(state))
    return grid, lt


# ---------------------------------------------------------------- backward (added capability)
# There is no reference counterpart: the eigendecomposition derivative is the textbook
# adjoint of T = Q diag(lam) Q^T restricted to the tridiagonal parameters (d, e).  It is
# pinned in tests/test_oracle.py by central finite differences of a LAPACK solve.

def eig_backward_dense(lam, Q, lam_bar, Q_bar):
    """(d_bar, e_bar) for cotangents of (lam, Q), by the dense formula
    T_bar = Q (diag(lam_bar) + F o (Q^T Q_bar)) Q^T, F_ij = 1/(lam_j - lam_i), F_ii = 0;
    d_bar = diag(T_bar), e_bar_k = T_bar[k, k+1] + T_bar[k+1, k]."""
    lam = np.asarray(lam, dtype=np.float64)
    Q = np.asarray(Q, dtype=np.float64)
    n = lam.shape[0]
    W = Q.T @ np.asarray(Q_bar, dtype=np.float64)
    D = lam[None, :] - lam[:, None]
    np.fill_diagonal(D, 1.0)
    F = 1.0 / D
    np.fill_diagonal(F, 0.0)
    M = F * W
    M[np.diag_indices(n)] = np.asarray(lam_bar, dtype=np.float64)
    Tb = Q @ M @ Q.T
    d_bar = np.diag(Tb).copy()
    e_bar = np.diag(Tb, 1) + np.diag(Tb, -1)
    return d_bar, e_bar


def eig_sorted_signfixed(d, e):
    """LAPACK eigendecomposition of T(d, e), eigenvalues ascending, each eigenvector's
    largest-magnitude entry made positive (a locally constant convention, so the loss
    L = w.lam + <G, Q> is differentiable under small perturbations)."""
    import scipy.linalg as sl
    lam, Q = sl.eigh_tridiagonal(np.asarray(d, dtype=np.float64), np.asarray(e, dtype=np.float64))
    idx = np.argmax(np.abs(Q), axis=0)
    sgn = np.sign(Q[idx, np.arange(Q.shape[1])])
    sgn[sgn == 0] = 1.0
    return lam, np.asfortranarray(Q * sgn[None, :])


def signfix_like(Q, Qref):
    """Flip columns of Q to agree in sign with Qref (same eigenvectors up to sign)."""
    s = np.sign(np.sum(Q * Qref, axis=0))
    s[s == 0] = 1.0
    return np.asfortranarray(Q * s[None, :])


def fd_directional(d, e, w, G, vd, ve, h: float = 1e-6):
    """Central difference of L(d, e) = w.lam + <G, Q> (sorted, sign-fixed LAPACK solve)
    along the direction (vd, ve): (L(x + h v) - L(x - h v)) / 2h, which equals
    grad L . v up to O(h^2)."""
    d = np.asarray(d, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    _, Q0 = eig_sorted_signfixed(d, e)

    def loss(dd, ee):
        lam, Q = eig_sorted_signfixed(dd, ee)
        Q = signfix_like(Q, Q0)
        return float(w @ lam + np.sum(G * Q))

    return (loss(d + h * vd, e + h * ve) - loss(d - h * vd, e - h * ve)) / (2 * h)


def fd_gradient(d, e, w, G, h: float = 1e-6, d_idx=None, e_idx=None):
    """Central differences of L(d, e) = w.lam + <G, Q> (sorted, sign-fixed LAPACK solve)
    on the chosen entries of d and e (all by default)."""
    d = np.asarray(d, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    _, Q0 = eig_sorted_signfixed(d, e)

    def loss(dd, ee):
        lam, Q = eig_sorted_signfixed(dd, ee)
        Q = signfix_like(Q, Q0)
        return float(w @ lam + np.sum(G * Q))

    d_idx = range(d.shape[0]) if d_idx is None else d_idx
    e_idx = range(e.shape[0]) if e_idx is None else e_idx
    gd = {}
    for k in d_idx:
        dp, dm = d.copy(), d.copy()
        dp[k] += h
        dm[k] -= h
        gd[k] = (loss(dp, e) - loss(dm, e)) / (2 * h)
    ge = {}
    for k in e_idx:
        ep, em = e.copy(), e.copy()
        ep[k] += h
        em[k] -= h
        ge[k] = (loss(d, ep) - loss(d, em)) / (2 * h)
    return gd, ge
