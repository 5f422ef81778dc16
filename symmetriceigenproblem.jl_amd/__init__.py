"""MI355X (gfx950) eigensolver hot path of SymmetricEigenProblem.jl, host side.

Python mirror of the reference's API over the C ABI of ``libsep_amd.so``
(``include/sep.h``).  Names follow the reference (``qr_algorithm!`` becomes
``qr_algorithm_``: the trailing underscore marks in-place mutation, Julia's ``!``).

* ``qr_algorithm_(H, Q, max_layers=128, tol=eps, maxiter=100n) -> (H, converged, iter)``
  mirrors src/bulge_chasing.jl:146 (same argument meaning, same return tuple, odd n
  raises ``AssertionError`` like the reference's ``@assert iseven(n)``).
* ``bulk_wave_order_2x2_rmul_(A, givens)`` and ``bulk_wave_order_2x2_rmul_parallel_``
  mirror src/wave.jl:62 and :36.
* ``eigen_decomp_(H)`` mirrors src/bulge_chasing.jl:290-299.
* ``qr_batched_``, ``eig_backward`` and ``eig_backward_batched`` are the added batched
  solver and backward pass.

Arrays: host arrays are numpy; device arrays are torch tensors on a ROCm device
(torch is used for device memory only).  Matrices are column-major: a numpy array must
be Fortran-ordered, a torch matrix must have strides (1, ld).  A (c,s) grid is the
memory of a Julia ``Matrix{Tuple{T,T}}`` of size (n-1) x k, i.e. an array of shape
(k, n-1, 2).

There is no CPU fallback: every entry point runs the HIP library and raises if it is
missing or if no gfx950 device is present.
"""
from __future__ import annotations

import ctypes as C
import logging
import os
import warnings
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import inputs  # noqa: F401  (seeded generators for the benchmark configurations)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsep_amd.so")

SEP_OK = 0
SEP_NOT_CONVERGED = 1
SEP_ERR_ARG = -1
SEP_ERR_ODD_N = -2
SEP_ERR_BAD_LAYERS = -3
SEP_ERR_HIP = -4
SEP_ERR_NOMEM = -5
SEP_ERR_UNSUPPORTED = -6

# Every symbol include/sep.h declares (checked by the CPU test suite).
EXPORTS = (
    "sep_create", "sep_destroy", "sep_last_error", "sep_version", "sep_set_stream",
    "sep_synchronize", "sep_set_timing", "sep_get_timing",
    "sep_bulk_rmul_f64", "sep_bulk_rmul_f32", "sep_bulk_rmul_f64_dev", "sep_bulk_rmul_f32_dev",
    "sep_bulk_rmul_f64_dev_noalloc",
    "sep_qr_algorithm_f64", "sep_qr_algorithm_f32", "sep_qr_algorithm_f64_dev",
    "sep_qr_algorithm_f32_dev",
    "sep_qr_batched_f64_dev", "sep_tune_batched_f64", "sep_set_batched_config",
    "sep_set_odd_order", "sep_set_chase_rotations",
    "sep_eig_backward_f64_dev", "sep_eig_backward_batched_f64_dev",
    "sep_sytrd_f64_dev", "sep_orgtr_f64_dev",
    "sep_sort_eigen_f64_dev", "sep_sort_eigen_f32_dev",
)

_lib: Optional[C.CDLL] = None


class SepError(RuntimeError):
    """A failing call into libsep_amd (message from sep_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libsep_amd error {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libsep_amd.so (built in-tree); raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run symmetriceigenproblem.jl_amd/build.py")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


_P = C.c_void_p
_I64 = C.c_int64
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)


def _declare(L: C.CDLL) -> None:
    L.sep_create.argtypes = [C.POINTER(_P), C.c_int]
    L.sep_destroy.argtypes = [_P]
    L.sep_last_error.restype = C.c_char_p
    L.sep_last_error.argtypes = []
    L.sep_version.restype = C.c_char_p
    L.sep_version.argtypes = []
    L.sep_set_stream.argtypes = [_P, _P]
    L.sep_synchronize.argtypes = [_P]
    L.sep_set_timing.argtypes = [_P, C.c_int]
    L.sep_get_timing.argtypes = [_P, _P, _P, _P]
    for suf, ct in (("f64", C.c_double), ("f32", C.c_float)):
        for dev in ("", "_dev"):
            f = getattr(L, f"sep_bulk_rmul_{suf}{dev}")
            f.argtypes = [_P, _I64, _I64, _P, _I64, _P, _I64]
            f = getattr(L, f"sep_qr_algorithm_{suf}{dev}")
            f.argtypes = [_P, _I64, _P, _P, _P, _I64, _I64, ct, _I64, _P, _P, _P]
    L.sep_bulk_rmul_f64_dev_noalloc.argtypes = [_P, _I64, _I64, _P, _I64, _P, _I64]
    L.sep_qr_batched_f64_dev.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _I64, _P, _P]
    L.sep_tune_batched_f64.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P]
    L.sep_set_batched_config.argtypes = [_P, C.c_int32, C.c_int32]
    L.sep_set_odd_order.argtypes = [_P, C.c_int32]
    L.sep_set_chase_rotations.argtypes = [_P, C.c_int32]
    L.sep_eig_backward_f64_dev.argtypes = [_P, _I64, _P, _P, _I64, _P, _P, _P, _P]
    L.sep_eig_backward_batched_f64_dev.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.sep_sytrd_f64_dev.argtypes = [_P, _I64, _P, _I64, _P, _P, _P]
    L.sep_orgtr_f64_dev.argtypes = [_P, _I64, _P, _I64, _P, _P, _I64]
    for suf in ("f64", "f32"):
        getattr(L, f"sep_sort_eigen_{suf}_dev").argtypes = [_P, _I64, _P, _P, _I64, _P, _P,
                                                             _I64, _P]
    for name in EXPORTS:
        if name not in ("sep_last_error", "sep_version"):
            getattr(L, name).restype = C.c_int


def last_error() -> str:
    return lib().sep_last_error().decode()


def version() -> str:
    return lib().sep_version().decode()


def _check(rc: int, allow_not_converged: bool = False) -> int:
    if rc == SEP_OK or (allow_not_converged and rc == SEP_NOT_CONVERGED):
        return rc
    msg = last_error()
    if rc == SEP_ERR_ODD_N:
        raise AssertionError(msg)
    if rc in (SEP_ERR_ARG, SEP_ERR_BAD_LAYERS):
        raise ValueError(msg)
    raise SepError(rc, msg)


def _init_torch_runtime() -> None:
    """Initialise torch's GPU runtime before the library's.  torch carries its own copy of
    the HIP runtime beside the one libsep_amd.so links; with both in one process, torch's
    must come up first (the other order leaves torch reporting no GPU)."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


class Handle:
    """Owns a sep_handle (device buffers, chase/apply streams, tuning table)."""

    def __init__(self, device: int = 0):
        _init_torch_runtime()
        self._h = _P()
        _check(lib().sep_create(C.byref(self._h), int(device)))
        self.device = device

    @property
    def ptr(self) -> int:
        return self._h.value

    def close(self) -> None:
        if self._h:
            lib().sep_destroy(self._h)
            self._h = _P()

    def synchronize(self) -> None:
        _check(lib().sep_synchronize(self._h))

    def set_stream(self, stream_ptr: Optional[int]) -> None:
        _check(lib().sep_set_stream(self._h, stream_ptr))

    def set_timing(self, enable: bool) -> None:
        """Bracket every bulk-apply kernel with HIP events (resets the counters)."""
        _check(lib().sep_set_timing(self._h, 1 if enable else 0))

    def set_batched_config(self, fused_depth: int = 0, rows_per_lane: int = 0) -> None:
        """Force the batched apply's fused depth / rows per lane (16/1, 8/1, 8/2, 4/1); 0/0 = tuned."""
        _check(lib().sep_set_batched_config(self._h, int(fused_depth), int(rows_per_lane)))

    def set_odd_order(self, allow: bool = True) -> None:
        """Accept odd orders in this handle's solves (an extension: the reference asserts
        iseven(n) as a limitation of its wave kernels, src/bulge_chasing.jl:155)."""
        _check(lib().sep_set_odd_order(self._h, int(bool(allow))))

    def set_chase_rotations(self, mode: str = "exact") -> None:
        """Rotation arithmetic of this handle's Float64 solves, single and batched:
        "exact" (the default, the reference's givensAlgorithm, bitwise-equal results) or
        "fast" (one reciprocal square root per chase rotation, results equal to rounding;
        no counterpart in the reference).  Float32 solves always use the exact rotation."""
        modes = {"exact": 0, "fast": 1}
        if mode not in modes:
            raise ValueError(f"mode must be one of {sorted(modes)}")
        _check(lib().sep_set_chase_rotations(self._h, modes[mode]))

    def get_timing(self) -> Tuple[float, float, int]:
        """(apply-kernel ms, preparation-kernels ms, apply launches) since the last reset."""
        a, p, n = C.c_double(), C.c_double(), C.c_int64()
        _check(lib().sep_get_timing(self._h, C.byref(a), C.byref(p), C.byref(n)))
        return a.value, p.value, n.value

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_default: Optional[Handle] = None


def default_handle() -> Handle:
    global _default
    if _default is None:
        _default = Handle(0)
    return _default


# ---------------------------------------------------------------- array helpers

def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _dtype_suffix(dt) -> str:
    s = str(dt)
    if s in ("float64", "torch.float64"):
        return "f64"
    if s in ("float32", "torch.float32"):
        return "f32"
    raise TypeError(f"unsupported dtype {dt} (Float64 / Float32 only)")


def _np_colmajor(A: np.ndarray) -> Tuple[int, int, int]:
    if A.ndim != 2 or not A.flags.f_contiguous:
        raise ValueError("matrix must be a 2-d Fortran-ordered (column-major) numpy array")
    return A.shape[0], A.shape[1], max(A.shape[0], 1)


def _require_device(name: str, t, dtype=None, device=None, min_len: Optional[int] = None,
                    vector: bool = False):
    """Raise ValueError unless t is a torch tensor on a GPU (the *_dev entry points take
    its data pointer), with the given dtype and device, stride 1 if a vector, and at least
    min_len elements."""
    if not _is_torch(t):
        raise ValueError(f"{name} must be a torch device tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} is on {t.device}; the device entry points need a GPU tensor")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} has dtype {t.dtype}, expected {dtype}")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if vector and (t.dim() != 1 or (t.numel() > 1 and t.stride(0) != 1)):
        raise ValueError(f"{name} must be a 1-d tensor with stride 1")
    if min_len is not None and t.numel() < min_len:
        raise ValueError(f"{name} has {t.numel()} elements, needs at least {min_len}")


def _torch_colmajor(A) -> Tuple[int, int, int]:
    if A.dim() != 2 or A.stride(0) != 1 or (A.shape[1] > 1 and A.stride(1) < A.shape[0]):
        raise ValueError("device matrix must be column-major: strides (1, ld) with ld >= rows")
    ld = A.stride(1) if A.shape[1] > 1 else max(A.shape[0], 1)
    return A.shape[0], A.shape[1], ld


def colmajor_empty(m: int, n: int, dtype, device="cuda"):
    """An uninitialised column-major m x n torch tensor (strides (1, m))."""
    import torch
    return torch.empty((n, m), dtype=dtype, device=device).t()


def to_device_colmajor(A: np.ndarray, device="cuda"):
    """Copy a host matrix to a column-major device tensor."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(A).T)).to(device).t()


def to_host_colmajor(A) -> np.ndarray:
    """Copy a column-major device tensor back to a Fortran-ordered numpy array."""
    return np.asfortranarray(A.t().contiguous().cpu().numpy().T)


# ---------------------------------------------------------------- reference API

@dataclass
class SymTridiagonal:
    """The two fields qr_algorithm! uses of LinearAlgebra.SymTridiagonal: dv (n), ev (n-1)."""

    dv: object
    ev: object

    def __post_init__(self):
        n = self.dv.shape[0]
        if self.ev.shape[0] != max(n - 1, 0):
            raise ValueError("ev must have length n-1")
        if str(self.dv.dtype) != str(self.ev.dtype):
            raise TypeError("dv and ev must share a dtype")

    @property
    def n(self) -> int:
        return int(self.dv.shape[0])

    def copy(self) -> "SymTridiagonal":
        return SymTridiagonal(self.dv.clone() if _is_torch(self.dv) else self.dv.copy(),
                              self.ev.clone() if _is_torch(self.ev) else self.ev.copy())

    def dense(self) -> np.ndarray:
        d = self.dv.cpu().numpy() if _is_torch(self.dv) else np.asarray(self.dv)
        e = self.ev.cpu().numpy() if _is_torch(self.ev) else np.asarray(self.ev)
        return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


def qr_algorithm_(H: SymTridiagonal, Q=None, max_layers: int = 128, tol=None,
                  maxiter: Optional[int] = None, handle: Optional[Handle] = None,
                  return_rotations: bool = False):
    """qr_algorithm!(H, Q, max_layers, tol, maxiter) -> (H, converged, iter).

    Mutates H.dv (eigenvalues, deflation order), H.ev and Q (Q <- Q * prod G^T) in place.
    Q may be None (eigenvalues only).  numpy inputs run through the host-pointer entry
    point, torch device tensors through the *_dev entry point."""
    n = H.n
    suf = _dtype_suffix(H.dv.dtype)
    npdt = np.float64 if suf == "f64" else np.float32
    if tol is None:
        tol = float(np.finfo(npdt).eps)
    if maxiter is None:
        maxiter = 100 * n
    dev = _is_torch(H.dv)
    if dev:
        _require_device("H.dv", H.dv, vector=True)
        _require_device("H.ev", H.ev, dtype=H.dv.dtype, device=H.dv.device, vector=True)
        if Q is not None:
            _require_device("Q", Q, dtype=H.dv.dtype, device=H.dv.device)
            qm, qn, ldq = _torch_colmajor(Q)
            if (qm, qn) != (n, n):
                raise ValueError("Q must be n x n with H's dtype")
    else:
        for a in (H.dv, H.ev):
            if not (isinstance(a, np.ndarray) and a.flags.c_contiguous and a.dtype == npdt):
                raise ValueError("dv/ev must be contiguous numpy arrays of one float dtype")
        if Q is not None:
            if _is_torch(Q):
                raise ValueError("Q must be a numpy array when H holds numpy arrays")
            qm, qn, ldq = _np_colmajor(Q)
            if (qm, qn) != (n, n) or Q.dtype != npdt:
                raise ValueError("Q must be n x n with H's dtype")
    h = handle or default_handle()
    conv = C.c_int32()
    it = C.c_int64()
    nrot = C.c_int64()
    if dev:
        qp = Q.data_ptr() if Q is not None else None
        if Q is None:
            ldq = max(n, 1)
        fn = getattr(lib(), f"sep_qr_algorithm_{suf}_dev")
        rc = fn(h.ptr, n, H.dv.data_ptr(), H.ev.data_ptr() if n > 1 else None, qp, ldq,
                int(max_layers), tol, int(maxiter), C.byref(conv), C.byref(it), C.byref(nrot))
    else:
        qp = Q.ctypes.data if Q is not None else None
        if Q is None:
            ldq = max(n, 1)
        fn = getattr(lib(), f"sep_qr_algorithm_{suf}")
        rc = fn(h.ptr, n, H.dv.ctypes.data, H.ev.ctypes.data if n > 1 else None, qp, ldq,
                int(max_layers), tol, int(maxiter), C.byref(conv), C.byref(it), C.byref(nrot))
    _check(rc, allow_not_converged=True)
    if return_rotations:
        return H, bool(conv.value), it.value, nrot.value
    return H, bool(conv.value), it.value


def bulk_wave_order_2x2_rmul_(A, givens, handle: Optional[Handle] = None):
    """bulk_wave_order_2x2_rmul!(A, givens): A <- A * prod_L prod_i G(i,L)^T in place.

    givens has shape (k, n-1, 2) (a Julia Matrix{Tuple{T,T}} of size (n-1) x k).  A fused
    block of rotations whose s are all zero is skipped, as in the reference (a rotation
    with s = 0 and c != 1 elsewhere is applied).  Unlike the reference there is no parity
    precondition on n or k."""
    if _is_torch(A):
        _require_device("A", A)
        _require_device("givens", givens, dtype=A.dtype, device=A.device)
        m, n, lda = _torch_colmajor(A)
        if not givens.is_contiguous():
            raise ValueError("givens must be contiguous")
        k = givens.shape[0] if givens.numel() else 0
        if givens.numel() and (givens.dim() != 3 or givens.shape[1] != n - 1
                               or givens.shape[2] != 2):
            raise ValueError("givens must have shape (k, n-1, 2)")
        suf = _dtype_suffix(A.dtype)
        h = handle or default_handle()
        fn = getattr(lib(), f"sep_bulk_rmul_{suf}_dev")
        _check(fn(h.ptr, m, n, A.data_ptr(), lda, givens.data_ptr(), k))
    else:
        if _is_torch(givens):
            raise ValueError("givens must be a numpy array when A is")
        m, n, lda = _np_colmajor(A)
        suf = _dtype_suffix(A.dtype)
        g = np.ascontiguousarray(givens, dtype=A.dtype)
        k = g.shape[0] if g.size else 0
        if g.size and (g.ndim != 3 or g.shape[1:] != (n - 1, 2)):
            raise ValueError("givens must have shape (k, n-1, 2)")
        h = handle or default_handle()
        fn = getattr(lib(), f"sep_bulk_rmul_{suf}")
        _check(fn(h.ptr, m, n, A.ctypes.data, lda, g.ctypes.data, k))
    return None


# The GPU entry point already spreads the rows of A over the whole device.
bulk_wave_order_2x2_rmul_parallel_ = bulk_wave_order_2x2_rmul_

_log = logging.getLogger("sep_amd")


def sort_eigen(lam, Q=None, handle: Optional[Handle] = None):
    """Eigenvalues in ascending order and Q's columns permuted to match (device tensors).

    Returns (lam_sorted, Q_sorted, perm) with lam_sorted[r] = lam[perm[r]] and column r of
    Q_sorted = column perm[r] of Q (new tensors; Q_sorted is None when Q is None).  Ties
    keep index order; NaNs go last."""
    import torch
    _require_device("lam", lam, vector=True)
    n = lam.shape[0]
    suf = _dtype_suffix(lam.dtype)
    ldq = max(n, 1)
    if Q is not None:
        _require_device("Q", Q, dtype=lam.dtype, device=lam.device)
        qm, qn, ldq = _torch_colmajor(Q)
        if (qm, qn) != (n, n):
            raise ValueError("Q must be n x n")
    h = handle or default_handle()
    lam_out = torch.empty_like(lam)
    perm = torch.empty(n, dtype=torch.int64, device=lam.device)
    Q_out = colmajor_empty(n, n, lam.dtype, lam.device) if Q is not None else None
    fn = getattr(lib(), f"sep_sort_eigen_{suf}_dev")
    _check(fn(h.ptr, n, lam.data_ptr(), Q.data_ptr() if Q is not None else None, ldq,
              lam_out.data_ptr(), Q_out.data_ptr() if Q is not None else None, max(n, 1),
              perm.data_ptr()))
    h.synchronize()
    return lam_out, Q_out, perm


def eigen_decomp_(H: SymTridiagonal, handle: Optional[Handle] = None, sort: bool = False,
                  max_layers: int = 128, maxiter: Optional[int] = None):
    """eigen_decomp!(H): Q = I, qr_algorithm!(H, Q), returns (eigenvalues, Q).

    Like the reference (src/bulge_chasing.jl:290-299) the eigenvalues are H.dv in
    deflation order and the solve's converged / iter are logged (logger "sep_amd", INFO);
    a solve that did not converge also raises a RuntimeWarning, since its Q lacks the
    unfinished block's rotations.  sort=True (device tensors) returns the eigenvalues in
    ascending order with Q's columns permuted to match (sort_eigen).  max_layers and
    maxiter pass through to qr_algorithm_ (the reference uses its defaults)."""
    n = H.n
    if _is_torch(H.dv):
        _require_device("H.dv", H.dv, vector=True)
        import torch
        Q = torch.eye(n, dtype=H.dv.dtype, device=H.dv.device).t()  # column-major view
    else:
        if sort:
            raise ValueError("sort=True needs device tensors")
        Q = np.asfortranarray(np.eye(n, dtype=H.dv.dtype))
    _, converged, it = qr_algorithm_(H, Q, max_layers, maxiter=maxiter, handle=handle)
    _log.info("done converged=%s iter=%d", converged, it)
    if not converged:
        warnings.warn(f"qr_algorithm! did not converge within maxiter (iter={it}); the "
                      "eigenvalues and Q are those of the last completed block",
                      RuntimeWarning, stacklevel=2)
    if sort:
        lam, Qs, _ = sort_eigen(H.dv, Q, handle=handle)
        return lam, Qs
    return H.dv, Q


# ---------------------------------------------------------------- batched / backward

def _host_i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _batch_layout(ns, off_d, off_q):
    """Host int64 copies of a packed batch layout and the buffer lengths it needs
    (dv/ev: max(off_d + n); Q: max(off_q + n^2)); raises ValueError on a malformed one."""
    ns, off_d, off_q = _host_i64(ns), _host_i64(off_d), _host_i64(off_q)
    if ns.ndim != 1 or off_d.shape != ns.shape or off_q.shape != ns.shape:
        raise ValueError("ns, off_d and off_q must be 1-d arrays of one length")
    if ns.size and (ns.min() < 0 or off_d.min() < 0 or off_q.min() < 0):
        raise ValueError("orders and offsets must be non-negative")
    need_d = int((off_d + ns).max()) if ns.size else 0
    need_q = int((off_q + ns * ns).max()) if ns.size else 0
    return ns, off_d, off_q, need_d, need_q


def qr_batched_(ns, off_d, off_q, dv, ev, Q, max_layers: int = 0, handle: Optional[Handle] = None):
    """Solve B problems (Float64, device tensors) with the round-based batched solver
    (rounds replayed from a captured hipGraph; SEP_BATCHED_GRAPH=0 at handle creation
    launches them one by one).

    Problem p: order ns[p] (even), dv/ev at off_d[p], Q (ns[p]^2, column-major) at off_q[p];
    Q may be None (eigenvalues only).
    Returns (converged[B] bool numpy, iters[B] int64 numpy)."""
    import torch
    ns, off_d, off_q, need_d, need_q = _batch_layout(ns, off_d, off_q)
    _require_device("dv", dv, dtype=torch.float64, min_len=need_d, vector=True)
    _require_device("ev", ev, dtype=torch.float64, device=dv.device, min_len=need_d - 1,
                    vector=True)
    if Q is not None:
        _require_device("Q", Q, dtype=torch.float64, device=dv.device, min_len=need_q)
        if not Q.is_contiguous():
            raise ValueError("Q must be a contiguous packed buffer")
    h = handle or default_handle()
    B = ns.shape[0]
    conv = torch.zeros(B, dtype=torch.int32, device=dv.device)
    iters = torch.zeros(B, dtype=torch.int64, device=dv.device)
    rc = lib().sep_qr_batched_f64_dev(h.ptr, B, ns.ctypes.data, off_d.ctypes.data,
                                      off_q.ctypes.data, dv.data_ptr(), ev.data_ptr(),
                                      Q.data_ptr() if Q is not None else None,
                                      int(max_layers), conv.data_ptr(),
                                      iters.data_ptr())
    _check(rc, allow_not_converged=True)
    torch.cuda.synchronize(dv.device)
    return conv.cpu().numpy().astype(bool), iters.cpu().numpy()


def tune_batched(ns, off_d, off_q, dv, ev, handle: Optional[Handle] = None) -> int:
    import torch
    ns, off_d, off_q, need_d, _ = _batch_layout(ns, off_d, off_q)
    _require_device("dv", dv, dtype=torch.float64, min_len=need_d, vector=True)
    _require_device("ev", ev, dtype=torch.float64, device=dv.device, min_len=need_d - 1,
                    vector=True)
    h = handle or default_handle()
    choice = C.c_int32(-1)
    _check(lib().sep_tune_batched_f64(h.ptr, ns.shape[0], ns.ctypes.data, off_d.ctypes.data,
                                      off_q.ctypes.data, dv.data_ptr(), ev.data_ptr(),
                                      C.byref(choice)))
    return choice.value


def eig_backward(lam, Q, lam_bar, Q_bar, handle: Optional[Handle] = None):
    """(d_bar, e_bar) for cotangents (lam_bar, Q_bar) of (lam, Q) = eigen(T(d, e)).

    All device tensors, Float64; Q and Q_bar column-major n x n with the same ld."""
    import torch
    _require_device("lam", lam, dtype=torch.float64, vector=True)
    n = lam.shape[0]
    _require_device("lam_bar", lam_bar, dtype=torch.float64, device=lam.device, min_len=n,
                    vector=True)
    _require_device("Q", Q, dtype=torch.float64, device=lam.device)
    _require_device("Q_bar", Q_bar, dtype=torch.float64, device=lam.device)
    qm, qn, ldq = _torch_colmajor(Q)
    bm, bn, ldb = _torch_colmajor(Q_bar)
    if (qm, qn) != (n, n) or (bm, bn) != (n, n):
        raise ValueError("Q and Q_bar must be n x n")
    if ldq != ldb:
        raise ValueError("Q and Q_bar must share a leading dimension")
    h = handle or default_handle()
    d_bar = torch.empty(n, dtype=torch.float64, device=lam.device)
    e_bar = torch.empty(max(n - 1, 1), dtype=torch.float64, device=lam.device)
    _check(lib().sep_eig_backward_f64_dev(h.ptr, n, lam.data_ptr(), Q.data_ptr(), ldq,
                                          lam_bar.data_ptr(), Q_bar.data_ptr(),
                                          d_bar.data_ptr(), e_bar.data_ptr()))
    return d_bar, e_bar[: max(n - 1, 0)]


def eig_backward_batched(ns, off_d, off_q, lam, Q, lam_bar, Q_bar, handle: Optional[Handle] = None):
    """Batched backward over packed problems (layout as qr_batched_)."""
    import torch
    ns, off_d, off_q, need_d, need_q = _batch_layout(ns, off_d, off_q)
    _require_device("lam", lam, dtype=torch.float64, min_len=need_d, vector=True)
    _require_device("lam_bar", lam_bar, dtype=torch.float64, device=lam.device, min_len=need_d,
                    vector=True)
    for name, t in (("Q", Q), ("Q_bar", Q_bar)):
        _require_device(name, t, dtype=torch.float64, device=lam.device, min_len=need_q)
        if not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous packed buffer")
    h = handle or default_handle()
    d_bar = torch.zeros_like(lam)
    e_bar = torch.zeros_like(lam)
    _check(lib().sep_eig_backward_batched_f64_dev(
        h.ptr, ns.shape[0], ns.ctypes.data, off_d.ctypes.data, off_q.ctypes.data,
        lam.data_ptr(), Q.data_ptr(), lam_bar.data_ptr(), Q_bar.data_ptr(),
        d_bar.data_ptr(), e_bar.data_ptr()))
    return d_bar, e_bar


# ---------------------------------------------------------------- dense input (no reference
# counterpart: the reference starts from a SymTridiagonal)

def sytrd_(A, handle: Optional[Handle] = None):
    """Reduce a dense symmetric A (Float64, column-major, lower triangle referenced) to
    tridiagonal T = Q^T A Q in place, LAPACK dsytrd conventions (uplo = 'L').

    Returns (d, e, tau); A then holds e on its subdiagonal and the reflectors below it.
    A torch device tensor runs in place on the device (d, e, tau are device tensors); a
    Fortran-ordered numpy array is copied to the device and back (numpy results)."""
    import torch
    host = not _is_torch(A)
    if not host:
        _require_device("A", A, dtype=torch.float64)
    elif A.dtype != np.float64:
        raise ValueError("A must be square Float64")
    h = handle or default_handle()
    Ad = to_device_colmajor(A) if host else A
    m, n, lda = _torch_colmajor(Ad)
    if m != n or Ad.dtype != torch.float64:
        raise ValueError("A must be square Float64")
    d = torch.empty(max(n, 1), dtype=torch.float64, device=Ad.device)
    e = torch.empty(max(n - 1, 1), dtype=torch.float64, device=Ad.device)
    tau = torch.empty(max(n - 1, 1), dtype=torch.float64, device=Ad.device)
    _check(lib().sep_sytrd_f64_dev(h.ptr, n, Ad.data_ptr(), lda, d.data_ptr(), e.data_ptr(),
                                   tau.data_ptr()))
    d, e, tau = d[:n], e[: max(n - 1, 0)], tau[: max(n - 1, 0)]
    if host:
        A[...] = to_host_colmajor(Ad)
        return d.cpu().numpy(), e.cpu().numpy(), tau.cpu().numpy()
    return d, e, tau


def orgtr(A, tau, handle: Optional[Handle] = None):
    """The orthogonal factor Q = H(0) ... H(n-2) of sytrd_'s output (LAPACK dorgtr, 'L')."""
    import torch
    host = not _is_torch(A)
    if not host:
        _require_device("A", A, dtype=torch.float64)
        _require_device("tau", tau, dtype=torch.float64, device=A.device,
                        min_len=max(A.shape[0] - 1, 0), vector=True)
    h = handle or default_handle()
    Ad = to_device_colmajor(A) if host else A
    td = torch.from_numpy(np.ascontiguousarray(tau)).to(Ad.device) if host else tau
    m, n, lda = _torch_colmajor(Ad)
    if m != n:
        raise ValueError("A must be square")
    Q = torch.empty((n, n), dtype=torch.float64, device=Ad.device).t()
    _check(lib().sep_orgtr_f64_dev(h.ptr, n, Ad.data_ptr(), lda,
                                   td.data_ptr() if n > 1 else None, Q.data_ptr(), max(n, 1)))
    return to_host_colmajor(Q) if host else Q


def eigh_dense_(A, max_layers: int = 128, handle: Optional[Handle] = None):
    """Eigen-decomposition of a dense symmetric Float64 A (lower triangle): sytrd_, orgtr,
    then qr_algorithm_ on T starting from that Q.  Returns (lam, Q, converged, iter) with
    lam in the solver's deflation order and A Q = Q diag(lam); A is overwritten."""
    d, e, tau = sytrd_(A, handle=handle)
    Q = orgtr(A, tau, handle=handle)
    H = SymTridiagonal(d, e)
    _, conv, it = qr_algorithm_(H, Q, max_layers, handle=handle)
    return H.dv, Q, conv, it

