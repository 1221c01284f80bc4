"""Pins the learned-preconditioner oracle (oracle/learned_oracle.py): its backward pass must
match central finite differences of its forward pass, the loss must be the A-norm error
fraction left by the coarse correction, and the untrained network must be coarse Jacobi."""
import numpy as np
import pytest
import scipy.sparse as sp

import learned_oracle as lo


def _problem(n=27, seed=0):
    rng = np.random.default_rng(seed)
    # SPD coarse operator: weighted graph Laplacian of a random geometric graph + shift
    pts = rng.random((n, 3))
    W = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            dist = np.linalg.norm(pts[i] - pts[j])
            if dist < 0.45:
                W[i, j] = W[j, i] = np.exp(rng.standard_normal())
    L = np.diag(W.sum(1)) - W
    Ac = L + np.diag(0.05 + rng.random(n))
    e_k = rng.standard_normal(n)
    rc = Ac @ e_k
    s = e_k @ rc
    return Ac, rc, s


def test_backward_matches_finite_differences():
    Ac, rc, s = _problem()
    theta = lo.init_params(3)
    rng = np.random.default_rng(4)
    theta = theta + 0.05 * rng.standard_normal(theta.size)   # nonzero output head too
    g, loss = lo.backward(theta, Ac, rc, s)
    idx = rng.choice(theta.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (lo.forward(tp, Ac, rc, s)[1] - lo.forward(tm, Ac, rc, s)[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_loss_is_error_fraction_and_exact_coarse_solve_minimises_it():
    Ac, rc, s = _problem(seed=5)
    theta = lo.init_params(6)
    e, loss, _ = lo.forward(theta, Ac, rc, s)
    e_k = np.linalg.solve(Ac, rc)
    frac = (e_k - e) @ Ac @ (e_k - e) / (e_k @ Ac @ e_k)
    assert loss == pytest.approx(frac, rel=1e-10)
    # the untrained network returns the coarse Jacobi estimate u = rc / diag(Ac)
    np.testing.assert_allclose(e, rc / np.diag(Ac), rtol=1e-12)


def test_scale_homogeneity():
    Ac, rc, s = _problem(seed=7)
    theta = lo.init_params(8) + 0.03
    e1, _ = lo.forward(theta, Ac, rc)
    e2, _ = lo.forward(theta, Ac, 1e-6 * rc)
    np.testing.assert_allclose(e2, 1e-6 * e1, rtol=1e-9)


def test_galerkin_operator():
    rng = np.random.default_rng(9)
    A = sp.random(40, 40, density=0.2, random_state=rng, format="csr")
    A = A + A.T
    agg = rng.integers(0, 7, 40)
    Ac = lo.galerkin(A, agg, 7)
    D = A.toarray()
    for I in range(7):
        for J in range(7):
            assert Ac[I, J] == pytest.approx(D[np.ix_(agg == I, agg == J)].sum(), abs=1e-12)
