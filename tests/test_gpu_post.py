"""GPU parity tests of the post-processing kernels (rs_heat_flux_device,
rs_stress_process_device) against the oracle's restatement of RSolverHeat::process and
RSolverStress::process (volume elements), and the equilibrium property of the node forces."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_heat_flux_matches_oracle(gpu, orc):
    rs = gpu
    n = 10
    xyz, tets = rs.kuhn_box(n, n, n)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    ne = tets.shape[0]
    k = rs.field_lognormal(3, ne, 0.0, 1.5)
    T = np.random.default_rng(4).random(xyz.shape[0])
    asm = rs.Assembly(xyz, tets, 1, fixed)
    dflux = rs.DeviceArray(3 * ne)
    asm.heat_flux_device(rs.DeviceArray.from_host(k), rs.DeviceArray.from_host(T), dflux)
    got = dflux.download().reshape(ne, 3)
    ref = orc.heat_flux(xyz, tets, k, T)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def _cantilever(rs, nx, ny, nz, L=4.0):
    xyz, tets = rs.kuhn_box(nx, ny, nz, L, 1.0, 1.0)
    ne = tets.shape[0]
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    E = rs.field_lognormal(3, ne, 0.0, 1.0)
    nu = np.full(ne, 0.3)
    tris = rs.plane_faces(xyz, tets, 0, L)
    trac = np.tile([0.0, 0.0, -1.0], (tris.shape[0], 1))
    return xyz, tets, fixed, E, nu, tris, trac


def _process(rs, asm, E, nu, u, nn, ne):
    dN, dS, dV, dF = rs.DeviceArray(ne), rs.DeviceArray(ne), rs.DeviceArray(ne), rs.DeviceArray(3 * nn)
    asm.stress_process_device(rs.DeviceArray.from_host(E), rs.DeviceArray.from_host(nu),
                              rs.DeviceArray.from_host(u.reshape(-1)), dN, dS, dV, dF)
    return dN.download(), dS.download(), dV.download(), dF.download().reshape(nn, 3)


def test_stress_process_matches_oracle(gpu, orc):
    rs = gpu
    xyz, tets, fixed, E, nu, _, _ = _cantilever(rs, 8, 4, 4)
    nn, ne = xyz.shape[0], tets.shape[0]
    u = np.random.default_rng(5).standard_normal((nn, 3)) * 1e-2
    asm = rs.Assembly(xyz, tets, 3, fixed)
    got = _process(rs, asm, E, nu, u, nn, ne)
    ref = orc.stress_process(xyz, tets, E, nu, u)
    for g, r, name in zip(got, ref, ("normal", "shear", "von Mises", "node forces")):
        assert np.abs(g - r).max() <= 1e-11 * np.abs(r).max(), name
    # deterministic: colour launches add the node forces in a fixed order
    again = _process(rs, asm, E, nu, u, nn, ne)
    assert np.array_equal(again[3], got[3])


def test_stress_node_forces_balance_the_loads(gpu):
    """After the solve, the node forces K u on the free DOFs equal the assembled loads (the
    traction), and the reactions on the clamped face balance the total traction."""
    rs = gpu
    xyz, tets, fixed, E, nu, tris, trac = _cantilever(rs, 16, 4, 4)
    nn, ne = xyz.shape[0], tets.shape[0]
    asm = rs.Assembly(xyz, tets, 3, fixed)
    b = asm.assemble_stress(E, nu, np.ones(ne), [0.0, 0.0, 0.0], np.zeros(3 * nn), tris, trac)
    conf = rs.MatrixSolverConf(n_outer_iterations=100000, solver_cvg_value=1e-11, stop_rule=rs.STOP_RELRES)
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray.from_host(np.zeros(asm.n_dof))
    info = rs.RMatrixSolver(conf).solve_device(asm.matrix, db, dx, rs.PRECOND_BLOCK_JACOBI, 3)
    assert info.converged == 1
    du = rs.DeviceArray.from_host(np.zeros(3 * nn))
    asm.dofs_to_nodes(dx, du)
    u = du.download().reshape(nn, 3)
    _, _, vm, F = _process(rs, asm, E, nu, u, nn, ne)
    book = asm.book()
    free = book != 0xFFFFFFFF
    Ff = np.empty(asm.n_dof)
    Ff[book[free]] = F.reshape(-1)[free]
    assert np.linalg.norm(Ff - b) <= 1e-8 * np.linalg.norm(b)
    total_traction = np.array([0.0, 0.0, -1.0]) * 1.0       # unit downward traction on the 1 x 1 end face
    reaction = F.reshape(-1)[~free].reshape(-1, 3).sum(axis=0)
    np.testing.assert_allclose(reaction, -total_traction, atol=1e-7)
    assert np.isfinite(vm).all() and vm.max() > 0
