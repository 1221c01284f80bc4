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


# ------------------------------------------------------------------------------ multilevel cycle
def _cheb_residual_poly(lam, k, x):
    """Chebyshev residual polynomial R_k(x) = T_k((theta - x)/delta) / T_k(theta/delta)."""
    hi, lo = 1.05 * lam, lam / 30.0
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    T = np.polynomial.chebyshev.Chebyshev.basis(k)
    return T((theta - x) / delta) / T(theta / delta)


@pytest.mark.parametrize("steps", [1, 2, 3, 4])
def test_chebyshev_smoother_closed_form(steps):
    # A = D: D^-1 A = I, so k steps from zero give z = (1 - R_k(1)) D^-1 r
    rng = np.random.default_rng(steps)
    d = 0.5 + rng.random(50)
    A = sp.diags(d).tocsr()
    r = rng.standard_normal(50)
    lam = 1.7
    z = lo.cheb_smooth(A, 1.0 / d, lo.cheb_coefs(lam), r, None, steps)
    np.testing.assert_allclose(z, (1.0 - _cheb_residual_poly(lam, steps, 1.0)) * r / d, rtol=1e-12)
    # error propagation on an eigenvector of D^-1 A with eigenvalue mu: e <- R_k(mu) e
    n = 40
    L = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    mu_k = 1.0 - np.cos(np.pi * 3 / (n + 1))            # eigenvalue of D^-1 L (D = 2 I)
    v = np.sin(np.pi * 3 * np.arange(1, n + 1) / (n + 1))
    x = lo.cheb_smooth(L, 0.5 * np.ones(n), lo.cheb_coefs(2.0), L @ v, None, steps)
    np.testing.assert_allclose(v - x, _cheb_residual_poly(2.0, steps, mu_k) * v, atol=1e-12)


def _heat_like(n=10, seed=3):
    """2-D five-point operator with random positive edge weights and a Dirichlet shift."""
    rng = np.random.default_rng(seed)
    N = n * n
    rows, cols, vals = [], [], []
    diag = np.full(N, 0.05)
    for i in range(n):
        for j in range(n):
            a = i * n + j
            for di, dj in ((1, 0), (0, 1)):
                if i + di < n and j + dj < n:
                    b = (i + di) * n + (j + dj)
                    w = np.exp(rng.standard_normal())
                    rows += [a, b]
                    cols += [b, a]
                    vals += [-w, -w]
                    diag[a] += w
                    diag[b] += w
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N)) + sp.diags(diag)
    m1 = np.array([(i // 2) * (n // 2) + (j // 2) for i in range(n) for j in range(n)])
    m2 = np.array([(i // 2) * (n // 4 + (n // 2) % 2) + (j // 2) for i in range(n // 2) for j in range(n // 2)])
    _, m2 = np.unique(m2, return_inverse=True)
    return A.tocsr(), [m1, m2]


def test_two_level_projection_property():
    A, maps = _heat_like()
    levels, Ps = lo.hierarchy(A, maps[:1])
    y = np.random.default_rng(0).standard_normal(levels[1].shape[0])
    r = A @ (Ps[0] @ y)
    # no smoothing: z = P A_c^-1 P^T r = P y exactly for r in A range(P)
    A0 = levels[0]
    Ac = levels[1]
    z = Ps[0] @ lo.coarse_exact(Ac, Ps[0].T @ r)
    np.testing.assert_allclose(z, Ps[0] @ y, rtol=1e-10, atol=1e-12)
    # and the Galerkin operator is the restriction of A to range(P)
    np.testing.assert_allclose(Ac.toarray(), (Ps[0].T @ A0 @ Ps[0]).toarray())


def test_symmetric_cycle_is_symmetric_positive_definite():
    A, maps = _heat_like()
    levels, Ps = lo.hierarchy(A, maps)
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    for pre in (1, 2):
        B = lambda r: lo.vcycle(levels, Ps, r, pre, pre, lo.coarse_exact)
        assert abs(x @ B(y) - y @ B(x)) < 1e-10 * abs(x @ B(x))
        assert x @ B(x) > 0 and y @ B(y) > 0
    # a plausible mistake (the post-smoother continuing the pre-smoother's step coefficients)
    # breaks the symmetry
    lv2, P2 = lo.hierarchy(A, maps[:1])
    dinv = 1.0 / A.diagonal()
    coefs = lo.cheb_coefs(lo.gershgorin(A))

    def bad(r):
        z = lo.cheb_smooth(A, dinv, coefs, r, None, 1)
        z = z + P2[0] @ lo.coarse_exact(lv2[1], P2[0].T @ (r - A @ z))
        return lo.cheb_smooth(A, dinv, coefs[1:], r, z, 1)
    assert abs(x @ bad(y) - y @ bad(x)) > 1e-4 * abs(x @ bad(x))


def test_cycle_preconditions_cg():
    import scipy.sparse.linalg as sla
    A, maps = _heat_like(16, 5)
    levels, Ps = lo.hierarchy(A, maps)
    b = np.random.default_rng(2).standard_normal(A.shape[0])

    def its(M):
        k = [0]
        sla.cg(A, b, rtol=1e-8, maxiter=2000, M=M, callback=lambda _: k.__setitem__(0, k[0] + 1))
        return k[0]
    jac = its(sla.LinearOperator(A.shape, matvec=lambda r: r / A.diagonal()))
    mg = its(sla.LinearOperator(A.shape, matvec=lambda r: lo.vcycle(levels, Ps, r, 1, 1, lo.coarse_exact)))
    assert mg < 0.5 * jac


def test_fcg_closed_forms():
    """The oracle's flexible CG: an exact preconditioner converges in one iteration; the identity
    gives plain CG, which terminates after as many iterations as A has distinct eigenvalues."""
    rng = np.random.default_rng(7)
    Q, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    A = Q @ np.diag(np.repeat([1.0, 3.0, 10.0], [10, 15, 15])) @ Q.T
    b = rng.standard_normal(40)
    x, its = lo.fcg(A, b, lambda r: np.linalg.solve(A, r), tol=1e-12)
    assert its == 1 and np.allclose(A @ x, b, atol=1e-10)
    x, its = lo.fcg(A, b, lambda r: r, tol=1e-10)
    assert its == 3 and np.allclose(A @ x, b, atol=1e-8)


def _cantilever_like(nx=8, ny=3, nz=3, seed=0):
    """A 3-DOF-per-node SPD operator on a box of nodes: three scalar Laplacians coupled by a
    small symmetric term, clamped (removed) at x = 0; coordinates and node book alongside."""
    g = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij"), -1)
    xyz = g.reshape(-1, 3).astype(np.float64) / nx
    nn = xyz.shape[0]
    fixed = xyz[:, 0] == 0
    book = np.full(3 * nn, 0xFFFFFFFF, dtype=np.uint64)
    free = np.nonzero(~fixed)[0]
    for k, i in enumerate(free):
        book[3 * i:3 * i + 3] = 3 * k + np.arange(3)
    return xyz, book, free


def test_component_bins_keep_components_apart():
    """Per-component bins: level 0 puts two DOFs in one aggregate exactly when they have the same
    component and their nodes share a bin (brute force over all pairs), and every level's
    aggregate holds a single component."""
    xyz, book, free = _cantilever_like()
    bins = (4, 2, 2)
    maps = lo.bins_hierarchy(xyz, book, 3, bins, 40, by_component=True)
    book1 = np.full(xyz.shape[0], 0xFFFFFFFF, dtype=np.uint64)   # the free nodes, one DOF each
    book1[free] = np.arange(free.size)
    node_maps = lo.bins_hierarchy(xyz, book1, 1, bins, 40 // 3)
    n_dof = 3 * free.size
    comp = np.arange(n_dof) % 3
    node = np.arange(n_dof) // 3
    m0, nb0 = maps[0], node_maps[0]
    same = m0[:, None] == m0[None, :]
    want = (comp[:, None] == comp[None, :]) & (nb0[node][:, None] == nb0[node][None, :])
    assert np.array_equal(same, want)
    # composed maps: still one component per aggregate on every level
    cur = np.arange(n_dof)
    for m in maps:
        cur = m[cur]
        for a in np.unique(cur):
            assert np.unique(comp[cur == a]).size == 1
    # the per-node aggregation of the same book mixes the components of a node
    mixed = lo.bins_hierarchy(xyz, book, 3, bins, 40)
    assert mixed[0][0] == mixed[0][1] == mixed[0][2]
