"""GPU parity tests of the scalar-transport path (rs_assemble_transport_device,
rs_transport_update_device: RSolverFluidHeat / RSolverFluidParticle) against the oracle, and its
GMRES + Jacobi solve against a direct solve of the oracle's system."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu

EOD = 0xFFFFFFFF


def problem(rs, n, speed, seed):
    xyz, tets = rs.kuhn_box(n, n, n)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(seed)
    ca = rng.uniform(0.5, 3.0, ne)
    k = rng.uniform(0.5, 2.0, ne)
    # a rotating flow about the box axis, sampled at the element centroids
    c = xyz[tets].mean(1) - 0.5
    vel = np.stack([-c[:, 1], c[:, 0], 0.3 * np.ones(ne)], 1) * speed
    q = rng.uniform(0.0, 5.0, nn)
    field = rng.uniform(0.0, 2.0, nn)
    fixed = (np.abs(xyz[:, 2]) < 1e-12) | (np.abs(xyz[:, 0] - 1) < 1e-12)
    return xyz, tets, ca, k, vel, q, field, fixed.astype(np.uint8)


@pytest.mark.parametrize("unsteady", [0, 1])
@pytest.mark.parametrize("speed", [0.1, 200.0])
def test_transport_matches_oracle(gpu, orc, unsteady, speed):
    rs = gpu
    xyz, tets, ca, k, vel, q, field, fixed = problem(rs, 6, speed, 40 + unsteady)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    conf = rs.TransportConf(0.6, 0.02, unsteady)
    db = rs.DeviceArray(asm.n_dof)
    asm.assemble_transport_device(conf, *(rs.DeviceArray.from_host(a.reshape(-1)) for a in (ca, k, vel, q, field)),
                                  db)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    A_ref, b_ref = orc.transport_assemble(xyz, tets, ca, k, vel, q, field, 0.6, 0.02, unsteady, book, n_dof)
    D = (asm.matrix.to_scipy() - A_ref).toarray()
    assert np.abs(D).max() <= 1e-12 * np.abs(A_ref.data).max()
    b = db.download()
    assert np.abs(b - b_ref).max() <= 1e-12 * np.abs(b_ref).max()
    x_ref = spla.spsolve(A_ref.tocsc(), b_ref)
    if speed < 1.0:
        # diffusion-dominated: GMRES(100) + Jacobi converges in a few cycles; in the
        # advection-dominated steady case it stagnates on the oracle as well
        sconf = rs.MatrixSolverConf(solver_type=rs.SOLVER_GMRES, n_outer_iterations=200, n_inner_iterations=100,
                                    solver_cvg_value=1e-12, stop_rule=rs.STOP_RELRES)
        dx = rs.DeviceArray.from_host(np.zeros(n_dof))
        info = rs.RMatrixSolver(sconf).solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
        assert info.converged == 1
        assert np.linalg.norm(dx.download() - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    else:
        dx = rs.DeviceArray.from_host(x_ref)
    dfield = rs.DeviceArray.from_host(field.copy())
    norms = asm.transport_update_device(dx, dfield)
    f_ref = field.copy()
    ref = orc.transport_update(book, dx.download(), f_ref)
    np.testing.assert_array_equal(dfield.download(), f_ref)
    np.testing.assert_allclose(norms, ref, rtol=1e-13)


@pytest.mark.parametrize("unsteady", [0, 1])
@pytest.mark.parametrize("speed", [0.1, 50.0])
def test_transport_row_gather_matches_colour_launches(gpu, unsteady, speed):
    """The row gather (default) and the colour launches assemble the same bits: the same element
    terms, summed per row in the same (colour) order."""
    rs = gpu
    xyz, tets, ca, k, vel, q, field, fixed = problem(rs, 7, speed, 60 + unsteady)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    conf = rs.TransportConf(0.6, 0.02, unsteady)
    args = [rs.DeviceArray.from_host(a.reshape(-1)) for a in (ca, k, vel, q, field)]
    out = {}
    for rows in (True, False):
        asm.set_method(rows)
        db = rs.DeviceArray(asm.n_dof)
        asm.assemble_transport_device(conf, *args, db)
        out[rows] = (asm.matrix.to_scipy(), db.download())
    A_rows, A_col = out[True][0], out[False][0]
    assert np.array_equal(A_rows.indptr, A_col.indptr) and np.array_equal(A_rows.indices, A_col.indices)
    assert np.array_equal(A_rows.data, A_col.data)
    assert np.array_equal(out[True][1], out[False][1])
