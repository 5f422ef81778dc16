"""Checker helpers for the dense -> tridiagonal entries (test infrastructure).

LAPACK's dsytrd (scipy.linalg.lapack, uplo = 'L') is the reference for d, e, tau and the
stored reflectors; scipy has no dorgtr, so the orthogonal factor is formed here from the
textbook definition Q = H(0) H(1) ... H(n-2), H(i) = I - tau_i v_i v_i^T, with
v_i[i+1] = 1 and v_i[i+2:] stored below the subdiagonal.  tests/test_tridiag_ref.py pins
this helper by Q^T A Q = T and Q^T Q = I on LAPACK's own output."""
import numpy as np
import scipy.linalg.lapack as lapack


def sym_matrix(n: int, seed: int, scale: float = 1.0) -> np.ndarray:
    r = np.random.default_rng(seed)
    X = r.standard_normal((n, n))
    return np.asfortranarray(scale * (X + X.T) / 2)


def lapack_sytrd(A: np.ndarray):
    """(c, d, e, tau) from LAPACK dsytrd with uplo = 'L'."""
    c, d, e, tau, info = lapack.dsytrd(np.asfortranarray(A), lower=1)
    assert info == 0
    return c, d, e, tau


def q_from_reflectors(c: np.ndarray, tau: np.ndarray) -> np.ndarray:
    n = c.shape[0]
    Q = np.eye(n)
    for i in range(n - 2, -1, -1):
        v = np.zeros(n)
        v[i + 1] = 1.0
        v[i + 2:] = c[i + 2:, i]
        Q -= tau[i] * np.outer(v, v @ Q)
    return np.asfortranarray(Q)


def tridiag(d: np.ndarray, e: np.ndarray) -> np.ndarray:
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
