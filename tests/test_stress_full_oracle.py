"""CPU pins of the full RSolverStress volume path in the oracle (or_stress_assemble_ex): the
thermal-expansion load, the consistent mass, the transient theta-scheme and the modal branch.

* free thermal expansion: with a uniform alpha dT, any E and nu per element and only rigid-body
  constraints (3-2-1), the solution is u = alpha dT (x - x0) (zero stress), exactly;
* the consistent mass equals the closed form rho V (1 + d_mn) / 20 on each component (and so
  carries the total mass), with no coupling between components;
* the stiffness of every branch equals the independent C_ijkl formulation, and transient
  Ae = M + alpha dt K;
* a transient step at rest in a rigid translation c (K c = 0) returns c;
* the modal lifting subtracts (K + M) u of the prescribed positions, as the reference does
  (src/rsolverstress.cpp:1719-1739).
"""
import numpy as np
import scipy.sparse.linalg as spla

from helpers import bary_gradients, dense_stress_K, kuhn_box_np

EOD = 0xFFFFFFFF


def dense_mass(xyz, tets, rho):
    nn = xyz.shape[0]
    M = np.zeros((3 * nn, 3 * nn))
    for e, t in enumerate(tets):
        _, V = bary_gradients(xyz[t])
        Me = rho[e] * V / 20.0 * (np.ones((4, 4)) + np.eye(4))
        for c in range(3):
            idx = 3 * t + c
            M[np.ix_(idx, idx)] += Me
    return M


def ordered(book):
    free = book != EOD
    return np.nonzero(free)[0][np.argsort(book[free])], np.nonzero(~free)[0]


def test_free_thermal_expansion(orc):
    xyz, tets = kuhn_box_np(3, 2, 2, 1.5, 1.0, 0.8)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(12)
    E, nu = rng.uniform(1.0, 5.0, ne), rng.uniform(0.1, 0.4, ne)
    a_dT = 2e-3
    alpha_t, dT = np.full(ne, 1e-5), np.full(ne, a_dT / 1e-5)
    fixed = np.zeros((nn, 3), np.uint8)
    o = int(np.argmin(np.linalg.norm(xyz, axis=1)))
    px = int(np.argmin(np.linalg.norm(xyz - [1.5, 0, 0], axis=1)))
    py = int(np.argmin(np.linalg.norm(xyz - [0, 1.0, 0], axis=1)))
    fixed[o, :] = 1
    fixed[px, 1:] = 1
    fixed[py, 2] = 1
    book, n_dof = orc.node_book(nn, tets, 3, fixed.reshape(-1))
    A, M, b = orc.stress_assemble_ex(xyz, tets, E, nu, np.ones(ne), [0, 0, 0], np.zeros(3 * nn), book, n_dof,
                                     alpha_t=alpha_t, dT=dT)
    assert M is None
    u = spla.spsolve(A.tocsc(), b)
    exact = (a_dT * xyz).reshape(-1)
    fi, _ = ordered(book)
    np.testing.assert_allclose(u, exact[fi], rtol=1e-10, atol=1e-14)
    assert np.abs(b).max() > 0


def test_modal_mass_stiffness_and_lifting(orc):
    xyz, tets = kuhn_box_np(3, 2, 2)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(13)
    E, nu, rho = rng.uniform(1.0, 5.0, ne), rng.uniform(0.1, 0.4, ne), rng.uniform(0.5, 2.0, ne)
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    u = np.where(fixed == 1, rng.standard_normal(3 * nn), 0.0)
    g = np.array([0.3, -0.2, -9.81])
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    A, M, b = orc.stress_assemble_ex(xyz, tets, E, nu, rho, g, u, book, n_dof, problem=orc.STRESS_MODAL)
    K = dense_stress_K(xyz, tets, E, nu)
    Md = dense_mass(xyz, tets, rho)
    fi, fx = ordered(book)
    np.testing.assert_allclose(A.toarray(), K[np.ix_(fi, fi)], rtol=0, atol=1e-12 * np.abs(K).max())
    np.testing.assert_allclose(M.toarray(), Md[np.ix_(fi, fi)], rtol=0, atol=1e-13 * np.abs(Md).max())
    f = np.zeros(3 * nn)
    for e, t in enumerate(tets):
        _, V = bary_gradients(xyz[t])
        for c in range(3):
            f[3 * t + c] += rho[e] * g[c] * V / 4.0
    expect = f[fi] - (K + Md)[np.ix_(fi, fx)] @ u[fx]
    np.testing.assert_allclose(b, expect, rtol=1e-10, atol=1e-12 * np.abs(expect).max())
    # total mass on every component, no coupling between components
    V_e = np.array([bary_gradients(xyz[t])[1] for t in tets])
    Mfull = dense_mass(xyz, tets, rho)
    for c in range(3):
        ones = np.zeros(3 * nn)
        ones[c::3] = 1.0
        assert abs(ones @ Mfull @ ones - rho @ V_e) <= 1e-13 * (rho @ V_e)
    Mo = M.toarray()
    comp = np.array([i % 3 for i in range(3 * nn)])[fi]
    assert np.all(Mo[comp[:, None] != comp[None, :]] == 0.0)


def test_transient_rigid_translation_stays(orc):
    xyz, tets = kuhn_box_np(2, 2, 3)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(14)
    E, nu, rho = rng.uniform(1.0, 5.0, ne), rng.uniform(0.1, 0.4, ne), rng.uniform(0.5, 2.0, ne)
    fixed = np.repeat(np.abs(xyz[:, 2]) < 1e-12, 3).astype(np.uint8)
    c = np.array([0.3, -0.7, 0.2])
    u = np.tile(c, nn)
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    for alpha, dt in ((1.0, 0.1), (0.5, 0.03)):
        A, _, b = orc.stress_assemble_ex(xyz, tets, E, nu, rho, [0, 0, 0], u, book, n_dof,
                                         problem=orc.STRESS_TRANSIENT, alpha=alpha, dt=dt)
        x = spla.spsolve(A.tocsc(), b)
        fi, _ = ordered(book)
        np.testing.assert_allclose(x, u[fi], rtol=1e-10)
        K = dense_stress_K(xyz, tets, E, nu)
        Md = dense_mass(xyz, tets, rho)
        np.testing.assert_allclose(A.toarray(), (Md + alpha * dt * K)[np.ix_(fi, fi)], rtol=0,
                                   atol=1e-12 * np.abs(Md).max())


def test_static_matches_the_previous_entry(orc):
    xyz, tets = kuhn_box_np(2, 2, 2)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(15)
    E, nu, rho = rng.uniform(1.0, 5.0, ne), rng.uniform(0.1, 0.4, ne), rng.uniform(0.5, 2.0, ne)
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    u = np.where(fixed == 1, rng.standard_normal(3 * nn), 0.0)
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    A0, b0 = orc.stress_assemble(xyz, tets, E, nu, rho, [0, 0, -1.0], u, None, None, book, n_dof)
    A1, _, b1 = orc.stress_assemble_ex(xyz, tets, E, nu, rho, [0, 0, -1.0], u, book, n_dof)
    # the oracle scatters under an OpenMP critical section, so the order of the element sums (and
    # the last bit) can differ from call to call
    np.testing.assert_allclose(A0.toarray(), A1.toarray(), rtol=0, atol=1e-14 * np.abs(A0.data).max())
    np.testing.assert_allclose(b0, b1, rtol=0, atol=1e-14 * np.abs(b0).max())
