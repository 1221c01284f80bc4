"""GPU parity test of the magnetostatics assembly (rs_assemble_magnetostatics_device) against the
oracle, and its solve with the (negative definite) system through Jacobi-CG."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_magnetostatics_matches_oracle_and_solves(gpu, orc):
    rs = gpu
    n = 6
    xyz, tets = rs.kuhn_box(n, n, n)
    nn = xyz.shape[0]
    x = xyz[:, 0]
    fixed_node = (np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)
    fixed = np.repeat(fixed_node, 3).astype(np.uint8)
    rng = np.random.default_rng(6)
    J = rng.standard_normal((nn, 3)) * 1e6
    H = np.where(fixed_node[:, None], rng.standard_normal((nn, 3)), 0.0)
    asm = rs.Assembly(xyz, tets, 3, fixed)
    db = rs.DeviceArray(asm.n_dof)
    asm.assemble_magnetostatics_device(rs.DeviceArray.from_host(J.reshape(-1)), rs.DeviceArray.from_host(H.reshape(-1)), db)
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    A_ref, b_ref = orc.magnetostatics_assemble(xyz, tets, J, H, book, n_dof)
    A = asm.matrix.to_scipy()
    A.sort_indices()
    A_ref.sort_indices()
    # the GPU pattern holds full 3x3 node blocks; the oracle stores what the reference adds
    D = (A - A_ref).toarray()
    assert np.abs(D).max() <= 1e-12 * np.abs(A_ref.data).max()
    b = db.download()
    assert np.abs(b - b_ref).max() <= 1e-12 * np.abs(b_ref).max()
    conf = rs.MatrixSolverConf(n_outer_iterations=10000, solver_cvg_value=1e-11, stop_rule=rs.STOP_RELRES)
    dx = rs.DeviceArray.from_host(np.zeros(asm.n_dof))
    info = rs.RMatrixSolver(conf).solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
    assert info.converged == 1
    import scipy.sparse.linalg as spla
    x_ref = spla.spsolve(A_ref.tocsc(), b_ref)
    xg = dx.download()
    assert np.linalg.norm(xg - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
