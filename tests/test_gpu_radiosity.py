"""GPU parity tests of the radiative-heat path (rs_assemble_radiosity_device,
rs_radiosity_process_device) against the oracle on a closed box enclosure, and its CG + Jacobi
solve, which balances the net heats (sum_i A_i q_i = 0 in a closed, reciprocal enclosure)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_radiosity_matches_oracle_and_balances(gpu, orc):
    import __graft_entry__ as ge
    rs = gpu
    rad = ge.load_module("radiation")
    xyz, tris, pp, pe, F, area, face = rad.box_enclosure(5)
    npch, ne = F.shape[0], tris.shape[0]
    rng = np.random.default_rng(8)
    em = rng.uniform(0.2, 0.95, ne)
    te = 400.0 + 100.0 * face.repeat(2) + rng.uniform(0, 50, ne)
    am = np.full(ne, 300.0)
    R = rs.Radiosity(xyz, tris, pp, pe)
    db = rs.DeviceArray(npch)
    R.assemble_device(*(rs.DeviceArray.from_host(a.reshape(-1)) for a in (F, em, te, am)), db)
    A_ref, b_ref = orc.radiosity_assemble(F, pp, pe, em, te, am)
    A = R.matrix.to_scipy().toarray()
    assert np.abs(A - A_ref).max() <= 1e-12 * np.abs(A_ref).max()
    b = db.download()
    assert np.abs(b - b_ref).max() <= 1e-12 * np.abs(b_ref).max()
    conf = rs.MatrixSolverConf(n_outer_iterations=1000, solver_cvg_value=1e-13, stop_rule=rs.STOP_RELRES)
    dq = rs.DeviceArray.from_host(np.zeros(npch))
    info = rs.RMatrixSolver(conf).solve_device(R.matrix, db, dq, rs.PRECOND_JACOBI, 1)
    assert info.converged == 1
    q = dq.download()
    q_ref = np.linalg.solve(A_ref, b_ref)
    assert np.linalg.norm(q - q_ref) <= 1e-9 * np.linalg.norm(q_ref)
    assert abs(area @ q) <= 1e-8 * (area @ np.abs(q))
    dh = rs.DeviceArray(ne)
    R.process_device(dq, dh)
    elem_area = np.full(ne, 0.5 / 25)
    h_ref = orc.radiosity_process(pp, pe, elem_area, q)
    np.testing.assert_allclose(dh.download(), h_ref, rtol=1e-13, atol=1e-13 * np.abs(h_ref).max())
