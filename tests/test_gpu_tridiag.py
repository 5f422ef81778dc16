"""GPU dense -> tridiagonal reduction and orthogonal factor (no counterpart in the
reference; they feed its qr_algorithm! from a dense symmetric matrix).

Against LAPACK dsytrd (uplo = 'L'): d and e to 1e-12 |A|, tau and the stored reflectors to
1e-10, the formed Q to 1e-9 sqrt(n) (the GPU sums its symmetric matrix-vector products in
another order; both are backward stable, and the forward differences grow with n).  Independently of LAPACK: Q^T Q = I and
Q^T A Q = T to 1e-14 n |A|, and the untouched strict upper triangle.  End to end
(sytrd, orgtr, qr_algorithm!): eigenvalues against LAPACK's eigvalsh to 1e-12 |A| and
|A Q - Q diag(lam)| <= 1e-13 n |A|."""
import numpy as np
import pytest

from tridiag_ref import lapack_sytrd, q_from_reflectors, sym_matrix, tridiag

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 4, 5, 63, 64, 65, 100, 128, 129, 300, 513]


@pytest.mark.parametrize("n", SIZES)
def test_sytrd_matches_lapack(sep, handle, n):
    A = sym_matrix(n, seed=n)
    c_ref, d_ref, e_ref, tau_ref = lapack_sytrd(A)
    G = A.copy(order="F")
    d, e, tau = sep.sytrd_(G, handle=handle)
    nA = max(np.linalg.norm(A), 1e-300)
    tol = 1e-12 * nA
    assert np.abs(d - d_ref).max() <= tol, (n, np.abs(d - d_ref).max() / nA)
    if n > 1:
        assert np.abs(e - e_ref).max() <= tol
        assert np.abs(tau - tau_ref).max() <= 1e-10
        low = np.tril(np.ones((n, n), bool), -2)
        assert np.abs(G[low] - c_ref[low]).max(initial=0.0) <= 1e-10
        assert np.array_equal(np.diag(G, -1), e)
    # the strict upper triangle is never referenced or written
    up = np.triu(np.ones((n, n), bool), 1)
    assert np.array_equal(G[up], A[up])
    # LAPACK-independent: Q^T A Q = T with the GPU's own factor
    Q = sep.orgtr(G, tau, handle=handle)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 1e-14 * max(n, 1) * 4
    assert np.linalg.norm(Q.T @ A @ Q - tridiag(d, e)) <= 1e-14 * max(n, 1) * 4 * nA
    if n > 1:
        # a forward comparison: two backward-stable reductions that sum in different
        # orders drift apart with n (measured 2.5e-10 at n = 513); a wrong reflector is O(1)
        Qr = q_from_reflectors(c_ref, tau_ref)
        assert np.linalg.norm(Q - Qr) <= 1e-9 * np.sqrt(n)


def test_strided_device_and_garbage_upper(sep, handle):
    """Device tensors with ld > n; the upper triangle holds NaN (never read)."""
    import torch
    n, ld = 200, 231
    A = sym_matrix(n, seed=5)
    _, d_ref, e_ref, _ = lapack_sytrd(A)
    big = torch.full((n, ld), float("nan"), dtype=torch.float64, device="cuda")
    L = np.tril(A)
    L[np.triu_indices(n, 1)] = np.nan
    big[:, :n] = torch.from_numpy(np.ascontiguousarray(L.T)).cuda()
    Ad = big.t()[:n, :]
    d, e, tau = sep.sytrd_(Ad, handle=handle)
    torch.cuda.synchronize()
    nA = np.linalg.norm(A)
    assert np.abs(d.cpu().numpy() - d_ref).max() <= 1e-12 * nA
    assert np.abs(e.cpu().numpy() - e_ref).max() <= 1e-12 * nA
    assert torch.isnan(big[:, n:]).all()                  # padding untouched
    Q = sep.orgtr(Ad, tau, handle=handle)
    Qh = sep.to_host_colmajor(Q)
    T = tridiag(d.cpu().numpy(), e.cpu().numpy())
    assert np.linalg.norm(Qh.T @ A @ Qh - T) <= 1e-13 * n * nA


@pytest.mark.parametrize("n", [2, 10, 64, 250, 600])
def test_dense_eigendecomposition(sep, handle, n):
    A = sym_matrix(n, seed=100 + n)
    G = A.copy(order="F")
    lam, Q, conv, it = sep.eigh_dense_(G, handle=handle)
    assert conv
    nA = np.linalg.norm(A)
    ref = np.linalg.eigvalsh(A)
    assert np.abs(np.sort(lam) - ref).max() <= 1e-12 * nA
    assert np.linalg.norm(A @ Q - Q * lam) <= 1e-13 * n * nA
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 1e-13 * n


def test_dense_n2048_on_device(sep, handle):
    """n = 2048 on the device: orthogonality, reduction residual and eigenvalues."""
    import torch
    n = 2048
    A = sym_matrix(n, seed=7)
    Ad = sep.to_device_colmajor(A)
    d, e, tau = sep.sytrd_(Ad, handle=handle)
    Q = sep.orgtr(Ad, tau, handle=handle)
    Qc = Q
    At = torch.from_numpy(A).cuda()
    T = torch.diag(d) + torch.diag(e, 1) + torch.diag(e, -1)
    nA = torch.linalg.norm(At).item()
    I = torch.eye(n, dtype=torch.float64, device="cuda")
    assert torch.linalg.norm(Qc.T @ Qc - I).item() <= 1e-14 * n * 4
    assert torch.linalg.norm(Qc.T @ At @ Qc - T).item() <= 1e-14 * n * 4 * nA
    ref = np.linalg.eigvalsh(A)
    got = np.linalg.eigvalsh(tridiag(d.cpu().numpy(), e.cpu().numpy()))
    assert np.abs(got - ref).max() <= 1e-12 * nA


def test_errors(sep, handle):
    import torch
    A = torch.zeros((4, 4), dtype=torch.float64, device="cuda")  # row-major view: rejected
    with pytest.raises(ValueError):
        sep.sytrd_(A, handle=handle)
    L = sep.lib()
    import ctypes as C
    d = torch.zeros(4, dtype=torch.float64, device="cuda")
    At = A.t()
    # ld < n
    assert L.sep_sytrd_f64_dev(handle.ptr, 4, At.data_ptr(), 3, d.data_ptr(), d.data_ptr(),
                               d.data_ptr()) == sep.SEP_ERR_ARG
    assert L.sep_orgtr_f64_dev(handle.ptr, 4, At.data_ptr(), 4, None, At.data_ptr(), 4) == \
        sep.SEP_ERR_ARG
    assert L.sep_sytrd_f64_dev(None, 4, At.data_ptr(), 4, d.data_ptr(), d.data_ptr(),
                               d.data_ptr()) == sep.SEP_ERR_ARG


@pytest.mark.parametrize("scale_exp", [600, -600])
def test_sytrd_extreme_magnitudes(sep, handle, scale_exp):
    """A scaled by 2^+-600, where the squares in the reflector norm leave the range and the
    scaled norm is taken: d, e against LAPACK dsytrd of the same matrix, 1e-12 relative."""
    n = 150
    A0 = sym_matrix(n, seed=31)
    A = A0 * 2.0 ** scale_exp
    _, d_ref, e_ref, _ = lapack_sytrd(A)
    G = A.copy(order="F")
    d, e, tau = sep.sytrd_(G, handle=handle)
    nA = np.linalg.norm(A0) * 2.0 ** scale_exp  # (the norm of A itself would underflow)
    assert np.all(np.isfinite(d)) and np.all(np.isfinite(e))
    assert np.abs(d - d_ref).max() <= 1e-12 * nA
    assert np.abs(e - e_ref).max() <= 1e-12 * nA


def test_dense_long_columns_on_device(sep, handle):
    """n = 8320: the first columns exceed the 8192 rows the column step keeps in registers
    and take its multi-pass form.  Checked on the device: Q^T Q = I and Q^T A Q = T."""
    import torch
    n = 8320
    gen = torch.Generator(device="cuda").manual_seed(17)
    X = torch.randn((n, n), generator=gen, dtype=torch.float64, device="cuda")
    A = (X + X.T) / 2
    del X
    Ad = A.clone().t().contiguous().t()  # column-major copy (A is symmetric)
    d, e, tau = sep.sytrd_(Ad, handle=handle)
    Q = sep.orgtr(Ad, tau, handle=handle)
    del Ad
    nA = torch.linalg.norm(A).item()
    I = torch.eye(n, dtype=torch.float64, device="cuda")
    assert torch.linalg.norm(Q.T @ Q - I).item() <= 4e-14 * n
    T = torch.diag(d) + torch.diag(e, 1) + torch.diag(e, -1)
    assert torch.linalg.norm(Q.T @ (A @ Q) - T).item() <= 4e-14 * n * nA


@pytest.mark.parametrize("scale_exp", [-1000, -1040])
def test_sytrd_near_underflow(sep, handle, scale_exp):
    """A scaled by 2^-1000 (normal entries whose reflector beta lies below dlarfg's safmin,
    2^-969) and 2^-1040 (subnormal entries, where 1 / (alpha - beta) overflows unless x and
    alpha are rescaled first): like LAPACK's dlarfg, the reflector is recomputed from x and
    alpha scaled by 2^969.  Finite factors, an orthogonal Q, and d, e, tau against LAPACK
    dsytrd of the same matrix: 1e-12 relative at 2^-1000; at 2^-1040 the subnormal inputs
    carry ~34 bits, so both reductions agree to 1e-5 (LAPACK itself moves by 2.4e-7
    against the unscaled reduction there)."""
    n = 80
    A0 = sym_matrix(n, seed=33)
    A = np.ldexp(A0, scale_exp)
    _, d_ref, e_ref, tau_ref = lapack_sytrd(A)
    G = A.copy(order="F")
    d, e, tau = sep.sytrd_(G, handle=handle)
    assert np.all(np.isfinite(d)) and np.all(np.isfinite(e)) and np.all(np.isfinite(tau))
    assert np.all(np.isfinite(G))
    tol = 1e-12 if scale_exp >= -1000 else 1e-5
    nA = np.ldexp(np.linalg.norm(A0), scale_exp)
    assert np.abs(d - d_ref).max() <= tol * nA
    assert np.abs(e - e_ref).max() <= tol * nA
    assert np.abs(tau - tau_ref).max() <= tol * 10
    Q = sep.orgtr(G, tau, handle=handle)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 1e-14 * n * 4
