"""CPU tests that pin the oracle (oracle/rs_oracle.c) before it is trusted as the checker.

Each pin is chosen so that a plausible mistake fails it: the reference's own known answers
(tests/tst_matrix_solver.cpp), direct sparse solves (scipy, a different method), CG's finite
termination on k distinct eigenvalues, closed-form element matrices computed by a different
formulation, conservation (K 1 = 0, sum f = Q V), exact patch solutions and rigid-body kernels.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import bary_gradients, dense_heat, dense_stress_K, kuhn_box_np, random_spd

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "tst_matrix_solver.json")


def golden_cases():
    with open(GOLDEN) as f:
        g = json.load(f)
    out = []
    for c in g["cases"]:
        r, cc, v = zip(*c["entries"])
        A = sp.csr_matrix((v, (r, cc)), shape=(c["n"], c["n"]))
        out.append((c["name"], A, np.array(c["b"]), np.array(c["x"]), g))
    return out


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_oracle_reference_known_answers(orc, case):
    """tests/tst_matrix_solver.cpp:33-77: CG, preconditioner NONE, cvg 1e-13, 1000 its."""
    _, A, b, x_exp, g = case
    conf = g["conf"]
    x, info = orc.solve(A, b, precond=orc.PRECOND_NONE, n_outer=conf["n_outer_iterations"],
                        n_inner=conf["n_inner_iterations"], cvg=conf["solver_cvg_value"])
    assert np.all(np.abs(x - x_exp) < g["tolerance"])
    assert info["iterations"] <= A.shape[0] + 1


def test_oracle_cg_finite_termination(orc):
    """CG on a matrix with 3 distinct eigenvalues converges in exactly 3 iterations."""
    d = np.array([1.0, 1.0, 2.0, 2.0, 5.0, 5.0, 5.0, 1.0])
    A = sp.diags(d).tocsr()
    b = np.arange(1.0, 9.0)
    x, info = orc.solve(A, b, precond=orc.PRECOND_NONE, cvg=1e-13)
    assert info["iterations"] == 3
    np.testing.assert_allclose(x, b / d, rtol=1e-12)
    # Jacobi turns it into the identity: one iteration
    x, info = orc.solve(A, b, precond=orc.PRECOND_JACOBI, cvg=1e-13)
    assert info["iterations"] == 1
    np.testing.assert_allclose(x, b / d, rtol=1e-12)


@pytest.mark.parametrize("precond,block", [(0, 1), (1, 1), (2, 3), (2, 4)])
def test_oracle_cg_matches_direct_solve(orc, precond, block):
    A = random_spd(120, 0.04, seed=3)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(120)
    x_direct = spla.spsolve(A.tocsc(), b)
    x, info = orc.solve(A, b, precond=precond, block_size=block, cvg=1e-14, n_outer=5000)
    assert np.linalg.norm(x - x_direct) / np.linalg.norm(x_direct) < 1e-9
    assert info["reason"] in (1, 2, 3)
    # equation scale of src/rmatrixsolver.cpp:166-171
    An = np.sqrt((A.data ** 2).sum())
    assert info["equation_scale"] == pytest.approx(1e9 / abs(np.linalg.norm(b) / An), rel=1e-12)


@pytest.mark.parametrize("precond", [0, 1, 2])
def test_oracle_gmres_matches_direct_solve(orc, precond):
    rng = np.random.default_rng(5)
    n = 90
    A = (sp.random(n, n, density=0.05, random_state=rng) + sp.eye(n) * 3.0).tocsr()  # non-symmetric
    b = rng.standard_normal(n)
    x_direct = spla.spsolve(A.tocsc(), b)
    x, info = orc.solve(A, b, precond=precond, block_size=3, solver=orc.SOLVER_GMRES,
                        n_outer=200, n_inner=30, cvg=1e-14)
    assert np.linalg.norm(x - x_direct) / np.linalg.norm(x_direct) < 1e-8
    # restart length >= n: one Krylov cycle solves in exact arithmetic
    x, info = orc.solve(A, b, precond=precond, block_size=3, solver=orc.SOLVER_GMRES,
                        n_outer=3, n_inner=n, cvg=1e-14)
    assert np.linalg.norm(x - x_direct) / np.linalg.norm(x_direct) < 1e-8


def test_oracle_preconditioners(orc):
    A = random_spd(12, 0.3, seed=7).tolil()
    A[5, 5] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    d, w = orc.precond_build(A, orc.PRECOND_JACOBI)
    diag = A.diagonal()
    exp = np.where(diag != 0, 1.0 / np.where(diag != 0, diag, 1), 0.0)
    assert w == 1
    np.testing.assert_allclose(d, exp, rtol=1e-15)
    # block-Jacobi: width 5 does not divide 12 -> 4; blocks are the inverses of diagonal blocks
    A2 = random_spd(12, 0.3, seed=8)
    data, w = orc.precond_build(A2, orc.PRECOND_BLOCK_JACOBI, 5)
    assert w == 4
    Ad = A2.toarray()
    for ib in range(3):
        blk = Ad[4 * ib:4 * ib + 4, 4 * ib:4 * ib + 4]
        np.testing.assert_allclose(data.reshape(12, 4)[4 * ib:4 * ib + 4], np.linalg.inv(blk), rtol=1e-10)
    x = np.arange(12.0)
    y = orc.precond_apply(orc.PRECOND_BLOCK_JACOBI, 12, 4, data, x)
    ref = np.concatenate([np.linalg.inv(Ad[4 * i:4 * i + 4, 4 * i:4 * i + 4]) @ x[4 * i:4 * i + 4] for i in range(3)])
    np.testing.assert_allclose(y, ref, rtol=1e-10)


def test_oracle_stopping_rule(orc):
    """RIterationInfo::hasConverged, src/riterationinfo.cpp:90-109."""
    cvg = 1e-8
    assert orc.has_converged(float("nan"), 1.0, cvg)
    assert orc.has_converged(1.0, float("inf"), cvg)
    assert orc.has_converged(0.0, 1.0, cvg)        # 0 is not normal
    assert orc.has_converged(1.0, 0.0, cvg)        # zero trend is not normal
    assert orc.has_converged(1e-9, -1e-3, cvg)
    assert orc.has_converged(1.0, 1e-300, cvg)     # subnormal trend
    assert orc.has_converged(1.0, 5e-29, cvg)      # |trend| < eps^2 (eps 1e-14)
    assert not orc.has_converged(1e-7, -1e-8, cvg)


def test_oracle_zero_rhs(orc):
    A = sp.diags([2.0, 3.0]).tocsr()
    x, info = orc.solve(A, np.zeros(2), precond=1)
    assert info["iterations"] == 1 and info["equation_scale"] == 1.0
    assert np.all(x == 0.0)


def test_oracle_node_book(orc, rs):
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.uint32)
    fixed = np.array([1, 0, 0, 1, 0, 0], np.uint8)        # node 5 is in no element
    book, n = orc.node_book(6, tets, 1, fixed)
    assert n == 3 and list(book) == [0xFFFFFFFF, 0, 1, 0xFFFFFFFF, 2, 0xFFFFFFFF]
    fixed3 = np.zeros(18, np.uint8)
    fixed3[[0, 1, 2, 4]] = 1
    book3, n3 = orc.node_book(6, tets, 3, fixed3)
    assert n3 == 18 - 4 - 3
    assert book3[3] == 0 and book3[5] == 1 and book3[6] == 2
    lib_book, lib_n = rs.node_book(6, tets, 3, fixed3)
    assert lib_n == n3 and np.array_equal(lib_book, book3)


def _heat_case(n, rs, sigma=0.7):
    xyz, tets = kuhn_box_np(n, n, n)
    ne = tets.shape[0]
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    t_node = np.where(np.abs(x - 1) < 1e-12, 1.0, 0.0)
    rng = np.random.default_rng(11)
    k = np.exp(sigma * rng.standard_normal(ne))
    q = rng.random(ne)
    return xyz, tets, fixed, t_node, k, q


@pytest.mark.parametrize("transient", [False, True])
def test_oracle_heat_assembly_vs_dense(orc, rs, transient):
    xyz, tets, fixed, t_node, k, q = _heat_case(3, rs)
    rng = np.random.default_rng(12)
    rhoc = 1.0 + rng.random(tets.shape[0])
    if transient:
        t_node = t_node + 0.1 * rng.random(t_node.size)  # previous-step field off the boundary
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    A, b = orc.heat_assemble(xyz, tets, k, q, t_node, book, n_dof, rhoc=rhoc, transient=transient,
                             alpha=1.0, dt=1e-3)
    Ad, bd, _, _ = dense_heat(xyz, tets, k, q, t_node, book, n_dof, rhoc=rhoc, transient=transient,
                              alpha=1.0, dt=1e-3)
    np.testing.assert_allclose(A.toarray(), Ad, rtol=1e-12, atol=1e-12 * abs(Ad).max())
    np.testing.assert_allclose(b, bd, rtol=1e-11, atol=1e-12 * abs(bd).max())
    assert abs(A - A.T).max() <= 1e-14 * abs(A).max()


def test_oracle_heat_conservation(orc):
    """No Dirichlet nodes: K 1 = 0 (constants in the kernel) and sum(f) = sum(Q V) = sum(Q)/ne."""
    xyz, tets = kuhn_box_np(3, 2, 2, 1.0, 2.0, 0.5)
    ne = tets.shape[0]
    rng = np.random.default_rng(13)
    k, q = np.exp(rng.standard_normal(ne)), rng.random(ne)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, np.zeros(xyz.shape[0], np.uint8))
    A, b = orc.heat_assemble(xyz, tets, k, q, np.zeros(xyz.shape[0]), book, n_dof)
    assert abs(A @ np.ones(n_dof)).max() < 1e-13 * abs(A).max()
    vols = np.array([bary_gradients(xyz[t])[1] for t in tets])
    assert vols.sum() == pytest.approx(1.0, rel=1e-13)
    assert b.sum() == pytest.approx((q * vols).sum(), rel=1e-13)


def test_oracle_heat_layered_patch(orc):
    """Q=0, T=0/1 at x=0/1, k constant on each x-layer of cells: the exact discrete solution is
    the 1-D harmonic profile T(x_i) = sum_{j<i} h/k_j / sum_j h/k_j (linear inside layers)."""
    n = 5
    xyz, tets = kuhn_box_np(n, 3, 3)
    cx = np.array([xyz[t][:, 0].mean() for t in tets])
    layer = np.floor(cx * n).astype(int)
    kl = np.array([0.3, 4.0, 1.0, 20.0, 0.7])
    k = kl[layer]
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    t_node = np.where(np.abs(x - 1) < 1e-12, 1.0, 0.0)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    A, b = orc.heat_assemble(xyz, tets, k, np.zeros(tets.shape[0]), t_node, book, n_dof)
    sol, info = orc.solve(A, b, precond=orc.PRECOND_JACOBI, cvg=1e-15, n_outer=2000,
                          stop_rule=orc.STOP_RELRES)
    R = np.concatenate([[0.0], np.cumsum(1.0 / kl)]) / np.sum(1.0 / kl)
    free = book != 0xFFFFFFFF
    ix = np.rint(x[free] * n).astype(int)
    np.testing.assert_allclose(sol[book[free]], R[ix], atol=1e-10)


def test_oracle_stress_assembly_vs_tensor(orc):
    xyz, tets = kuhn_box_np(2, 2, 2, 2.0, 1.0, 1.0)
    ne, nn = tets.shape[0], xyz.shape[0]
    rng = np.random.default_rng(14)
    E, nu, rho = np.exp(rng.standard_normal(ne)), np.full(ne, 0.3), np.ones(ne)
    book, n_dof = orc.node_book(nn, tets, 3, np.zeros(3 * nn, np.uint8))
    A, b = orc.stress_assemble(xyz, tets, E, nu, rho, [0, 0, 0], np.zeros(3 * nn), None, None,
                               book, n_dof)
    Kd = dense_stress_K(xyz, tets, E, nu)
    np.testing.assert_allclose(A.toarray(), Kd, rtol=1e-11, atol=1e-12 * abs(Kd).max())
    # six rigid-body modes in the kernel
    X = xyz
    modes = [np.tile(e, nn) for e in np.eye(3)]
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1.0
        modes.append(np.cross(w, X).ravel())
    for u in modes:
        assert abs(A @ u).max() < 1e-12 * abs(A).max() * max(1.0, abs(u).max())


def test_oracle_stress_uniform_strain_patch(orc):
    """Homogeneous material, u = eps.x prescribed on the boundary: the interior reproduces the
    linear field exactly (with a heterogeneous E the stress would jump across elements)."""
    n = 3
    xyz, tets = kuhn_box_np(n, n, n)
    nn, ne = xyz.shape[0], tets.shape[0]
    on_bnd = np.any((np.abs(xyz) < 1e-12) | (np.abs(xyz - 1) < 1e-12), axis=1)
    eps = np.array([[1e-3, 2e-4, -3e-4], [2e-4, -5e-4, 1e-4], [-3e-4, 1e-4, 7e-4]])
    u_exact = xyz @ eps.T
    fixed = np.repeat(on_bnd, 3).astype(np.uint8)
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    E = np.full(ne, 2.5)
    A, b = orc.stress_assemble(xyz, tets, E, np.full(ne, 0.3), np.ones(ne), [0, 0, 0],
                               u_exact.ravel(), None, None, book, n_dof)
    sol, _ = orc.solve(A, b, precond=orc.PRECOND_JACOBI, cvg=1e-15, n_outer=5000,
                       stop_rule=orc.STOP_RELRES)
    free = book != 0xFFFFFFFF
    np.testing.assert_allclose(sol[book[free]], u_exact.ravel()[free], atol=1e-12)


def test_oracle_traction_total(orc):
    """Traction t on the x=L face: total load in b equals t * face area (nothing fixed)."""
    xyz, tets = kuhn_box_np(3, 2, 2, 3.0, 1.0, 1.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    faces = []
    for t in tets:
        for f in ([1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]):
            tri = t[f]
            if np.all(np.abs(xyz[tri, 0] - 3.0) < 1e-12):
                faces.append(tri)
    faces = np.array(faces, np.uint32)
    assert faces.shape[0] == 2 * 2 * 2
    trac = np.tile([0.0, 0.0, -1.0], (faces.shape[0], 1))
    book, n_dof = orc.node_book(nn, tets, 3, np.zeros(3 * nn, np.uint8))
    _, b = orc.stress_assemble(xyz, tets, np.ones(ne), np.full(ne, 0.3), np.ones(ne), [0, 0, 0],
                               np.zeros(3 * nn), faces, trac, book, n_dof)
    assert b[2::3].sum() == pytest.approx(-1.0, rel=1e-13)
    assert abs(b[0::3]).max() == 0.0
