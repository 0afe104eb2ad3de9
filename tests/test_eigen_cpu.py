"""CPU checks of the engine's host eigensolvers (qkrylov_amd_tridiag_eigen,
qkrylov_amd_tridiag_lowest, qkrylov_amd_hermitian_eigen): the Lanczos energy,
the Ritz coefficients, the FTLM weights and Davidson's projected problem come
from them. They are pinned against closed forms (the tridiagonal Toeplitz
spectrum), LAPACK through scipy/numpy, and invariants (orthonormality, the
first-row weights summing to one, residuals); no device is involved."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from qkrylov_amd import _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def tri_eig(alpha, beta, want):
    L = _lib.load()
    n = len(alpha)
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta if len(beta) else np.zeros(1), dtype=np.float64)
    w = np.zeros(max(n, 1))
    z = np.zeros(max(n * n if want == 2 else n, 1))
    _lib.check(L.qkrylov_amd_tridiag_eigen(n, _p(a), _p(b), _p(w), _p(z), want), "tridiag_eigen")
    if want == 2:
        return w[:n], z[:n * n].reshape(n, n)
    return w[:n], z[:n]


def tri_lowest(alpha, beta):
    L = _lib.load()
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta if len(beta) else np.zeros(1), dtype=np.float64)
    out = np.zeros(1)
    _lib.check(L.qkrylov_amd_tridiag_lowest(len(a), _p(a), _p(b), _p(out)), "tridiag_lowest")
    return out[0]


def herm_eig(A):
    L = _lib.load()
    n = A.shape[0]
    a = np.ascontiguousarray(A.astype(np.complex128)).view(np.float64)
    w = np.zeros(max(n, 1))
    v = np.zeros(2 * n * n if n else 2)
    _lib.check(L.qkrylov_amd_hermitian_eigen(n, _p(a), _p(w), _p(v)), "hermitian_eigen")
    return w[:n], v[:2 * n * n].view(np.complex128).reshape(n, n)


def lanczos_tridiag(H, v0, m):
    """Plain three-term Lanczos without reorthogonalisation (loses
    orthogonality, so its tridiagonal carries near-degenerate ghost pairs)."""
    v = v0 / np.linalg.norm(v0)
    vp = np.zeros_like(v)
    al, be = [], []
    b = 0.0
    for k in range(m):
        w = H @ v - b * vp
        a = v @ w
        w -= a * v
        al.append(a)
        b = np.linalg.norm(w)
        if k + 1 < m:
            be.append(b)
        vp, v = v, w / b
    return np.array(al), np.array(be)


def test_toeplitz_closed_form():
    n, a, b = 37, 0.3, -1.7
    w, z = tri_eig(np.full(n, a), np.full(n - 1, b), 2)
    k = np.arange(1, n + 1)
    exact = np.sort(a + 2 * b * np.cos(k * np.pi / (n + 1)))
    assert np.max(np.abs(w - exact)) < 1e-13
    # eigenvectors: sin(j k pi / (n+1)) up to sign and order
    j = np.arange(1, n + 1)
    for i in range(n):
        kk = np.argmin(np.abs(a + 2 * b * np.cos(k * np.pi / (n + 1)) - w[i])) + 1
        ex = np.sin(j * kk * np.pi / (n + 1))
        ex /= np.linalg.norm(ex)
        assert abs(abs(ex @ z[:, i]) - 1.0) < 1e-12
    assert abs(tri_lowest(np.full(n, a), np.full(n - 1, b)) - exact[0]) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 3, 10, 100, 300])
def test_random_tridiagonal_against_lapack(n):
    rng = np.random.default_rng(n)
    al = rng.standard_normal(n) * 3
    be = rng.standard_normal(n - 1)
    w, z = tri_eig(al, be, 2)
    wr, zr = sla.eigh_tridiagonal(al, be)
    scale = max(1.0, np.max(np.abs(wr)))
    assert np.max(np.abs(w - wr)) < 1e-13 * scale * max(1, np.sqrt(n))
    assert np.allclose(z.T @ z, np.eye(n), atol=1e-12)
    T = np.diag(al) + np.diag(be, 1) + np.diag(be, -1)
    assert np.max(np.abs(T @ z - z * w)) < 1e-12 * scale
    w1, f = tri_eig(al, be, 1)
    assert np.array_equal(w1, w)
    assert np.allclose(f, z[0], atol=1e-14)
    assert abs(tri_lowest(al, be) - wr[0]) < 4e-15 * scale * max(1, n)


def test_lanczos_ghost_pairs_keep_weights():
    """A 100-step Lanczos tridiagonal of a 60-dimensional matrix has more
    eigenvalues than the matrix: copies of converged ones. The first-row
    weights must still sum to one and reproduce the moments <v0|T^k|v0>."""
    rng = np.random.default_rng(7)
    D = 60
    Q, _ = np.linalg.qr(rng.standard_normal((D, D)))
    H = Q @ np.diag(np.linspace(-3, 2, D)) @ Q.T
    al, be = lanczos_tridiag(H, rng.standard_normal(D), 100)
    w, f = tri_eig(al, be, 1)
    assert abs(np.sum(f ** 2) - 1.0) < 1e-12
    T = np.diag(al) + np.diag(be, 1) + np.diag(be, -1)
    e0 = np.zeros(len(al))
    e0[0] = 1
    for k in range(1, 6):
        mom_ref = e0 @ np.linalg.matrix_power(T, k) @ e0
        assert abs(np.sum(f ** 2 * w ** k) - mom_ref) < 1e-11 * 3 ** k
    # exact spectrum is present: the lowest Ritz value is the lowest eigenvalue
    assert abs(w[0] + 3.0) < 1e-12
    assert abs(tri_lowest(al, be) + 3.0) < 1e-12


def test_decoupled_and_graded():
    # zero couplings: the diagonal itself, with unit vectors
    d = np.array([3.0, -1.0, 2.0, -1.0])
    w, z = tri_eig(d, np.zeros(3), 2)
    assert np.array_equal(w, np.sort(d))
    assert np.allclose(np.abs(z), np.abs(z).round())
    # strongly graded matrix
    al = 10.0 ** -np.arange(12, dtype=float)
    be = 10.0 ** -(np.arange(11) + 0.5)
    w, _ = tri_eig(al, be, 0)
    wr = sla.eigh_tridiagonal(al, be, eigvals_only=True)
    assert np.max(np.abs(w - wr)) < 1e-15


def test_wilkinson_clusters_orthogonal():
    m = 10
    al = np.abs(np.arange(-m, m + 1, dtype=float))
    be = np.ones(2 * m)
    w, z = tri_eig(al, be, 2)
    wr = sla.eigh_tridiagonal(al, be, eigvals_only=True)
    assert np.max(np.abs(w - wr)) < 1e-13
    assert np.allclose(z.T @ z, np.eye(2 * m + 1), atol=1e-12)


@pytest.mark.parametrize("n", [1, 2, 5, 20, 40])
def test_hermitian_against_lapack(n):
    rng = np.random.default_rng(100 + n)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = (X + X.conj().T) / 2
    w, V = herm_eig(A)
    wr = np.linalg.eigvalsh(A)
    assert np.max(np.abs(w - wr)) < 1e-13 * max(1, np.max(np.abs(wr)))
    assert np.allclose(V.conj().T @ V, np.eye(n), atol=1e-12)
    assert np.max(np.abs(A @ V - V * w)) < 1e-12 * max(1, np.max(np.abs(wr)))


def test_hermitian_degenerate_and_real():
    rng = np.random.default_rng(3)
    n = 12
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u /= np.linalg.norm(u)
    A = 2.0 * np.eye(n) - 5.0 * np.outer(u, u.conj())   # eigenvalues -3 (once), 2 (n-1 times)
    w, V = herm_eig(A)
    assert abs(w[0] + 3.0) < 1e-13 and np.max(np.abs(w[1:] - 2.0)) < 1e-13
    assert abs(abs(np.vdot(u, V[:, 0])) - 1.0) < 1e-12
    assert np.allclose(V.conj().T @ V, np.eye(n), atol=1e-12)
    # a real symmetric input gives the real spectrum
    S = rng.standard_normal((n, n))
    S = S + S.T
    w, V = herm_eig(S.astype(complex))
    assert np.max(np.abs(w - np.linalg.eigvalsh(S))) < 1e-13


def test_invalid_arguments():
    L = _lib.load()
    z = np.zeros(4)
    assert L.qkrylov_amd_tridiag_eigen(-1, _p(z), _p(z), _p(z), _p(z), 0) == -1
    assert L.qkrylov_amd_tridiag_eigen(2, _p(z), _p(z), _p(z), None, 2) == -1
    assert L.qkrylov_amd_tridiag_lowest(0, _p(z), _p(z), _p(z)) == -1
    assert L.qkrylov_amd_hermitian_eigen(2, None, _p(z), _p(z)) == -1
