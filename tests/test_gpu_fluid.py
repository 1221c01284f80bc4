"""GPU parity tests of the fluid path (rs_assemble_fluid_device, rs_fluid_update_device) against
the oracle, and a steady channel flow driven to convergence by Picard steps on the GPU
(assembly, GMRES with Jacobi, write-back) against the same steps on the oracle with direct
solves."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu

EOD = 0xFFFFFFFF


def channel(rs, nx, ny, nz, lx=1.0):
    xyz, tets = rs.kuhn_box(nx, ny, nz, lx, 1.0, 1.0)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    near = lambda a, v: np.abs(a - v) < 1e-12
    wall = near(y, 0) | near(y, 1) | near(z, 0) | near(z, 1)
    fixed = np.zeros((xyz.shape[0], 4), np.uint8)
    fixed[wall | near(x, 0), :3] = 1
    fixed[near(x, lx), 3] = 1
    inflow = near(x, 0) & ~wall
    return xyz, tets, fixed.reshape(-1), inflow


@pytest.mark.parametrize("unsteady", [0, 1])
@pytest.mark.parametrize("speed", [1e-3, 20.0])
def test_fluid_assembly_matches_oracle(gpu, orc, unsteady, speed):
    rs = gpu
    xyz, tets, fixed, _ = channel(rs, 6, 4, 3)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(11 + unsteady)
    rho = rng.uniform(0.5, 2.0, ne)
    mu = rng.uniform(0.01, 0.1, ne)
    grav = np.tile([0.0, 0.0, -9.81], (ne, 1)) + rng.standard_normal((ne, 3))
    state = np.zeros((nn, 4))
    state[:, :3] = rng.standard_normal((nn, 3)) * speed
    state[:, 3] = rng.standard_normal(nn)
    v_old = state[:, :3] + rng.standard_normal((nn, 3)) * 0.1 * speed
    acc = rng.standard_normal((nn, 3)) * speed
    conf = rs.FluidConf(1.7, 0.6, 0.05, unsteady)
    asm = rs.Assembly(xyz, tets, 4, fixed)
    db = rs.DeviceArray(asm.n_dof)
    dev = [rs.DeviceArray.from_host(np.ascontiguousarray(a).reshape(-1)) for a in (rho, mu, grav, state, v_old, acc)]
    asm.assemble_fluid_device(conf, *dev, db)
    book, n_dof = orc.node_book(nn, tets, 4, fixed)
    assert n_dof == asm.n_dof
    np.testing.assert_array_equal(asm.book(), book)
    A_ref, b_ref = orc.fluid_assemble(xyz, tets, rho, mu, grav, state, v_old, acc, 1.7, 0.6, 0.05, unsteady, book,
                                      n_dof)
    A = asm.matrix.to_scipy()
    D = (A - A_ref).toarray()
    assert np.abs(D).max() <= 1e-12 * np.abs(A_ref.data).max()
    b = db.download()
    assert np.abs(b - b_ref).max() <= 1e-12 * np.abs(b_ref).max()


@pytest.mark.parametrize("unsteady", [0, 1])
def test_fluid_update_matches_oracle(gpu, orc, unsteady):
    rs = gpu
    xyz, tets, fixed, _ = channel(rs, 5, 3, 3)
    nn = xyz.shape[0]
    asm = rs.Assembly(xyz, tets, 4, fixed)
    book = asm.book()
    rng = np.random.default_rng(21)
    x = rng.standard_normal(asm.n_dof)
    state = rng.standard_normal((nn, 4))
    v_old = rng.standard_normal((nn, 3))
    acc = np.zeros((nn, 3))
    ds, dvo, dacc = (rs.DeviceArray.from_host(a.reshape(-1)) for a in (state, v_old, acc))
    norms = asm.fluid_update_device(rs.DeviceArray.from_host(x), unsteady, 0.25, ds, dvo, dacc)
    ref = orc.fluid_update(book, x, unsteady, 0.25, state, v_old, acc)
    np.testing.assert_array_equal(ds.download(), state.reshape(-1))
    np.testing.assert_array_equal(dvo.download(), v_old.reshape(-1))
    if unsteady:
        np.testing.assert_array_equal(dacc.download(), acc.reshape(-1))
    np.testing.assert_allclose(norms, ref, rtol=1e-13)


def test_fluid_channel_picard(gpu, orc):
    """Steady duct flow (parabolic inflow, walls, outflow pressure 0), rho = 1, mu = 0.05: eight
    Picard steps on the GPU (GMRES + Jacobi to relres 1e-12) and on the oracle (direct solves)
    reach the same state; the residual falls by more than 1e4."""
    rs = gpu
    xyz, tets, fixed, inflow = channel(rs, 8, 4, 4, lx=2.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    y, z = xyz[:, 1], xyz[:, 2]
    state = np.zeros((nn, 4))
    state[inflow, 0] = 16 * y[inflow] * (1 - y[inflow]) * z[inflow] * (1 - z[inflow])
    v_old = state[:, :3].copy()
    acc = np.zeros((nn, 3))
    rho, mu, grav = np.ones(ne), np.full(ne, 0.05), np.zeros((ne, 3))
    asm = rs.Assembly(xyz, tets, 4, fixed)
    book = asm.book()
    conf = rs.FluidConf(1.0, 1.0, 1.0, 0)
    drho, dmu, dg = (rs.DeviceArray.from_host(a.reshape(-1)) for a in (rho, mu, grav))
    ds, dvo, dacc = (rs.DeviceArray.from_host(a.reshape(-1).copy()) for a in (state, v_old, acc))
    db, dx = rs.DeviceArray(asm.n_dof), rs.DeviceArray(asm.n_dof)
    sconf = rs.MatrixSolverConf(solver_type=rs.SOLVER_GMRES, n_outer_iterations=1000, n_inner_iterations=100,
                                solver_cvg_value=1e-12, stop_rule=rs.STOP_RELRES)
    solver = rs.RMatrixSolver(sconf)
    st, vo = state.copy(), v_old.copy()
    res = []
    for it in range(8):
        asm.assemble_fluid_device(conf, drho, dmu, dg, ds, dvo, dacc, db)
        res.append(np.linalg.norm(db.download()))
        dx.upload(np.zeros(asm.n_dof))
        info = solver.solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
        assert info.converged == 1, info
        asm.fluid_update_device(dx, False, 1.0, ds, dvo, dacc)
        A_ref, b_ref = orc.fluid_assemble(xyz, tets, rho, mu, grav, st, vo, acc, 1.0, 1.0, 1.0, 0, book, asm.n_dof)
        orc.fluid_update(book, spla.spsolve(A_ref.tocsc(), b_ref), 0, 1.0, st, vo, acc)
    assert res[-1] < 1e-4 * res[0]
    sg = ds.download().reshape(nn, 4)
    assert np.abs(sg - st).max() <= 1e-8 * np.abs(st).max()


def test_fluid_duct_mixed_gmres(gpu, orc):
    """One Picard system of the duct flow (16 x 4 x 4 cells, mu 0.02) solved by restarted GMRES +
    Jacobi to relres 1e-8 in FP64 and in mixed precision (FP32 basis, SpMV and preconditioner):
    both within 1e-6 of the oracle's direct solution, the mixed one within 10% of the FP64
    one's inner iterations."""
    rs = gpu
    xyz, tets, fixed, inflow = channel(rs, 16, 4, 4, lx=4.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    y, z = xyz[:, 1], xyz[:, 2]
    state = np.zeros((nn, 4))
    state[inflow, 0] = 16 * y[inflow] * (1 - y[inflow]) * z[inflow] * (1 - z[inflow])
    rho, mu, grav = np.ones(ne), np.full(ne, 0.02), np.zeros((ne, 3))
    asm = rs.Assembly(xyz, tets, 4, fixed)
    conf = rs.FluidConf(1.0, 1.0, 1.0, 0)
    dev = [rs.DeviceArray.from_host(np.ascontiguousarray(a).reshape(-1)) for a in
           (rho, mu, grav, state, state[:, :3].copy(), np.zeros((nn, 3)))]
    db = rs.DeviceArray(asm.n_dof)
    asm.assemble_fluid_device(conf, *dev, db)
    A_ref, b_ref = orc.fluid_assemble(xyz, tets, rho, mu, grav, state, state[:, :3].copy(), np.zeros((nn, 3)), 1.0,
                                      1.0, 1.0, 0, asm.book(), asm.n_dof)
    x_star = spla.spsolve(A_ref.tocsc(), b_ref)
    its, xs = {}, {}
    for name, prec in (("fp64", rs.PRECISION_FP64), ("mixed", rs.PRECISION_MIXED)):
        sconf = rs.MatrixSolverConf(solver_type=rs.SOLVER_GMRES, n_outer_iterations=100, n_inner_iterations=100,
                                    solver_cvg_value=1e-8, stop_rule=rs.STOP_RELRES, precision=prec)
        dx = rs.DeviceArray.from_host(np.zeros(asm.n_dof))
        info = rs.RMatrixSolver(sconf).solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
        assert info.converged == 1, info
        assert info.true_relres < 1e-8
        xs[name], its[name] = dx.download(), info.inner_iterations
        assert np.linalg.norm(xs[name] - x_star) / np.linalg.norm(x_star) <= 1e-6
    print(f"duct 16x4x4 GMRES inner iterations: fp64 {its['fp64']}, mixed {its['mixed']}", flush=True)
    assert its["mixed"] <= 1.1 * its["fp64"] + 2
