"""CPU pins of the REigenValueSolver restatement (oracle.eigen_solve): Krylov exactness on a
problem with three distinct eigenvalues, the Lanczos coefficients in closed form from the same
C-library random start, and Rayleigh's fixed point being an eigenvalue."""
import ctypes as C

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

LIBC = C.CDLL(None)
LIBC.rand.restype = C.c_int


def _rand_unit(n, seed):
    LIBC.srand(C.c_uint(seed))
    RAND_MAX = 2147483647
    v = np.array([LIBC.rand() / RAND_MAX for _ in range(n)], np.float64)
    return v / np.linalg.norm(v)


def _tridiag(n):
    K = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr() * 10.0
    M = sp.diags([np.ones(n - 1), 4 * np.ones(n), np.ones(n - 1)], [-1, 0, 1]).tocsr() / 6.0
    return M, K


def test_arnoldi_exact_on_three_distinct_eigenvalues(orc):
    """K^-1 M with three distinct eigenvalues: a 3-vector Krylov space holds them exactly, so
    Arnoldi + the Gram-Schmidt QR iteration return them (post-processed to 1/x, sorted)."""
    n = 30
    lam = np.repeat([2.0, 5.0, 11.0], 10)
    K = sp.diags(lam).tocsr()
    M = sp.identity(n, format="csr")
    LIBC.srand(1)
    d, ev, _ = orc.eigen_solve(M, K, 1, n_eigen=3, n_iter=2000, eig_cvg=1e-14, cvg=1e-14)
    np.testing.assert_allclose(d, [2.0, 5.0, 11.0], rtol=1e-9)
    assert ev.shape == (3, n)
    # every returned vector lies in one eigenspace: K v = lambda M v for one of the three
    for row in ev:
        r = [np.linalg.norm(K @ row - l * row) / np.linalg.norm(row) for l in (2.0, 5.0, 11.0)]
        assert min(r) < 1e-8


def test_lanczos_coefficients_closed_form(orc):
    """The reference's QL step does not rotate, so Lanczos returns |1/alpha_i| sorted, with
    alpha_0 = v0.K^-1 M v0 and alpha_1 from the next Lanczos vector; the start vector is the
    C library's rand() stream after srand(1)."""
    n = 25
    M, K = _tridiag(n)
    Kd, Md = K.toarray(), M.toarray()
    v0 = _rand_unit(n, 1)
    w0 = np.linalg.solve(Kd, Md @ v0)
    a0 = w0 @ v0
    w0 = w0 - a0 * v0
    e1 = np.linalg.norm(w0)
    v1 = w0 / e1
    w1 = np.linalg.solve(Kd, Md @ v1 - e1 * (Kd @ v0))
    a1 = w1 @ v1
    LIBC.srand(1)
    d, ev, _ = orc.eigen_solve(M, K, 0, n_eigen=2, cvg=1e-15)
    assert ev is None
    np.testing.assert_allclose(d, np.sort([abs(1 / a0), abs(1 / a1)]), rtol=1e-9)
    # a rotating QL would have returned the tridiagonal's eigenvalues instead
    T = np.array([[a0, e1], [e1, a1]])
    assert not np.allclose(np.sort(np.abs(1 / np.linalg.eigvalsh(T))), d, rtol=1e-6)


def test_rayleigh_fixed_point_is_an_eigenvalue(orc):
    n = 20
    M, K = _tridiag(n)
    lam = sla.eigh(K.toarray(), M.toarray(), eigvals_only=True)
    LIBC.srand(3)
    d, ev, _ = orc.eigen_solve(M, K, 2, n_iter=500, eig_cvg=1e-14, cvg=1e-15)
    mu = d[0]
    gaps = np.abs(1 / lam - mu) / mu
    # within 1e-5 of one eigenvalue of K^-1 M, while its neighbours are several percent away
    assert gaps.min() < 1e-5 and np.sort(gaps)[1] > 1e-2
    b = ev[0]
    # b is the matching eigenvector: K^-1 M b = mu b
    r = np.linalg.solve(K.toarray(), M.toarray() @ b) - mu * b
    assert np.linalg.norm(r) / np.linalg.norm(b) / mu < 1e-2
