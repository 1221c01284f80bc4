"""GPU edge cases of the widened entry points: systems without unknowns, a single radiating
patch (closed form), a pure-diffusion transport patch test solved by GMRES on the device, and
assemblies of the wrong kind rejected with an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIGMA = 5.670367e-8


def test_fluid_without_unknowns(gpu):
    rs = gpu
    xyz, tets = rs.kuhn_box(2, 2, 2)
    nn, ne = xyz.shape[0], tets.shape[0]
    asm = rs.Assembly(xyz, tets, 4, np.ones(4 * nn, np.uint8))
    assert asm.n_dof == 0 and asm.nnz == 0
    state = np.arange(4 * nn, dtype=np.float64)
    dev = [rs.DeviceArray.from_host(a) for a in (np.ones(ne), np.ones(ne), np.zeros(3 * ne), state,
                                                  np.zeros(3 * nn), np.zeros(3 * nn))]
    db = rs.DeviceArray(1)
    asm.assemble_fluid_device(rs.FluidConf(1.0, 1.0, 1.0, 0), *dev, db)
    norms = asm.fluid_update_device(rs.DeviceArray(1), False, 1.0, dev[3], dev[4], dev[5])
    np.testing.assert_array_equal(dev[3].download(), state)
    assert norms[0] == norms[1]


def test_single_gray_patch(gpu):
    rs = gpu
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float64)
    tris = np.array([[0, 1, 2], [1, 3, 2]], np.uint32)
    R = rs.Radiosity(xyz, tris, np.array([0, 2], np.uint32), np.array([0, 1], np.uint32))
    E, T, Ta = 0.6, 800.0, 300.0
    db = rs.DeviceArray(1)
    R.assemble_device(*(rs.DeviceArray.from_host(a) for a in (np.zeros(1), np.full(2, E), np.full(2, T),
                                                               np.full(2, Ta))), db)
    conf = rs.MatrixSolverConf(n_outer_iterations=10, solver_cvg_value=1e-14, stop_rule=rs.STOP_RELRES)
    dq = rs.DeviceArray.from_host(np.zeros(1))
    rs.RMatrixSolver(conf).solve_device(R.matrix, db, dq, rs.PRECOND_JACOBI, 1)
    q = E * SIGMA * (Ta ** 4 - T ** 4)
    np.testing.assert_allclose(dq.download(), [q], rtol=1e-12)
    dh = rs.DeviceArray(2)
    R.process_device(dq, dh)
    np.testing.assert_allclose(dh.download(), [q / 2, q / 2], rtol=1e-12)


def test_transport_pure_diffusion_linear_field(gpu):
    rs = gpu
    n = 8
    xyz, tets = rs.kuhn_box(n, n, n)
    nn, ne = xyz.shape[0], tets.shape[0]
    T = 1.0 + xyz @ np.array([0.5, -1.0, 2.0])
    fixed = ((np.abs(xyz) < 1e-12) | (np.abs(xyz - 1) < 1e-12)).any(1)
    asm = rs.Assembly(xyz, tets, 1, fixed.astype(np.uint8))
    db = rs.DeviceArray(asm.n_dof)
    field = np.where(fixed, T, 0.0)
    asm.assemble_transport_device(rs.TransportConf(1.0, 1.0, 0),
                                  *(rs.DeviceArray.from_host(a) for a in (np.full(ne, 1.5), np.full(ne, 0.2),
                                                                          np.zeros(3 * ne), np.zeros(nn), field)),
                                  db)
    conf = rs.MatrixSolverConf(solver_type=rs.SOLVER_GMRES, n_outer_iterations=100, n_inner_iterations=100,
                               solver_cvg_value=1e-13, stop_rule=rs.STOP_RELRES)
    dx = rs.DeviceArray.from_host(np.zeros(asm.n_dof))
    info = rs.RMatrixSolver(conf).solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
    assert info.converged == 1
    dfield = rs.DeviceArray.from_host(field.copy())
    asm.transport_update_device(dx, dfield)
    # a linear field is exact for linear elements; the clamp at zero does not bite (T > 0 inside)
    np.testing.assert_allclose(dfield.download(), T, rtol=1e-10, atol=1e-10)


def test_wrong_assembly_kind_is_rejected(gpu):
    rs = gpu
    xyz, tets = rs.kuhn_box(2, 2, 2)
    nn, ne = xyz.shape[0], tets.shape[0]
    a1 = rs.Assembly(xyz, tets, 1, np.zeros(nn, np.uint8))
    a3 = rs.Assembly(xyz, tets, 3, np.zeros(3 * nn, np.uint8))
    d = rs.DeviceArray(4 * nn)
    with pytest.raises(rs.RSError, match="4 dofs per node"):
        a3.assemble_fluid_device(rs.FluidConf(1.0, 1.0, 1.0, 0), d, d, d, d, d, d, d)
    with pytest.raises(rs.RSError, match="1 dof per node"):
        a3.assemble_transport_device(rs.TransportConf(1.0, 1.0, 0), d, d, d, d, d, d)
    with pytest.raises(rs.RSError, match="3 dofs per node"):
        a1.assemble_magnetostatics_device(d, d, d)
    a4 = rs.Assembly(xyz, tets, 4, np.zeros(4 * nn, np.uint8))
    with pytest.raises(rs.RSError, match="stream velocity"):
        a4.assemble_fluid_device(rs.FluidConf(0.0, 1.0, 1.0, 0), d, d, d, d, d, d, d)


def test_stress_ex_and_rotation_errors(gpu):
    rs = gpu
    xyz, tets = rs.kuhn_box(2, 2, 2)
    nn, ne = xyz.shape[0], tets.shape[0]
    a3 = rs.Assembly(xyz, tets, 3, np.zeros(3 * nn, np.uint8))
    ones = np.ones(ne)
    with pytest.raises(rs.RSError, match="unknown problem"):
        a3.assemble_stress_ex(rs.StressConf(7, 1.0, 1.0, (0.0, 0.0, 0.0)), ones, np.full(ne, 0.3), ones,
                              np.zeros(3 * nn))
    with pytest.raises(rs.RSError, match="null"):
        a3.assemble_stress_ex(rs.StressConf(0, 1.0, 1.0, (0.0, 0.0, 0.0)), ones, np.full(ne, 0.3), ones,
                              np.zeros(3 * nn), thermal_expansion=ones)
    with pytest.raises(rs.RSError, match="no modal assembly"):
        a3.mass_matrix()
    R = np.tile(np.eye(3), (nn, 1, 1))
    R[1] = 0.0
    with pytest.raises(rs.RSError, match="singular"):
        a3.set_rotations(R, np.ones(nn, np.uint8))
    a1 = rs.Assembly(xyz, tets, 1, np.zeros(nn, np.uint8))
    with pytest.raises(rs.RSError, match="3 dofs per node"):
        a1.set_rotations(np.tile(np.eye(3), (nn, 1, 1)), np.ones(nn, np.uint8))
