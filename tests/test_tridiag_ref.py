"""Pins the dense -> tridiagonal checker (tests/tridiag_ref.py) on the CPU: LAPACK's
reflectors, formed into Q by the helper, satisfy Q^T A Q = T and Q^T Q = I, and T has A's
eigenvalues; a planted error in the helper (a dropped reflector) is detected."""
import numpy as np
import pytest

from tridiag_ref import lapack_sytrd, q_from_reflectors, sym_matrix, tridiag


@pytest.mark.parametrize("n", [2, 3, 7, 64, 129])
def test_reflector_product_reconstructs(n):
    A = sym_matrix(n, seed=n)
    c, d, e, tau = lapack_sytrd(A)
    Q = q_from_reflectors(c, tau)
    T = tridiag(d, e)
    nA = np.linalg.norm(A)
    assert np.linalg.norm(Q.T @ A @ Q - T) <= 1e-14 * n * nA
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 1e-14 * n
    assert np.allclose(np.linalg.eigvalsh(T), np.linalg.eigvalsh(A), rtol=0, atol=1e-13 * n * nA)
    # e is also left on the subdiagonal of the output
    assert np.array_equal(np.diag(c, -1), e)


def test_planted_error_is_detected():
    n = 40
    A = sym_matrix(n, seed=1)
    c, d, e, tau = lapack_sytrd(A)
    bad = tau.copy()
    bad[n // 2] = 0.0  # drop one reflector
    Q = q_from_reflectors(c, bad)
    assert np.linalg.norm(Q.T @ A @ Q - tridiag(d, e)) > 1e-3 * np.linalg.norm(A)
