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


def test_magnetostatics_row_gather_matches_colour_launches(gpu):
    """The row gather (default) and the colour launches assemble the same bits."""
    rs = gpu
    n = 7
    xyz, tets = rs.kuhn_box(n, n, n)
    nn = xyz.shape[0]
    x = xyz[:, 0]
    fixed_node = (np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)
    fixed = np.repeat(fixed_node, 3).astype(np.uint8)
    rng = np.random.default_rng(8)
    J = rng.standard_normal((nn, 3)) * 1e6
    H = rng.standard_normal((nn, 3))
    asm = rs.Assembly(xyz, tets, 3, fixed)
    dJ, dH = rs.DeviceArray.from_host(J.reshape(-1)), rs.DeviceArray.from_host(H.reshape(-1))
    out = {}
    for rows in (True, False):
        asm.set_method(rows)
        db = rs.DeviceArray(asm.n_dof)
        asm.assemble_magnetostatics_device(dJ, dH, db)
        out[rows] = (asm.matrix.to_scipy(), db.download())
    A_rows, A_col = out[True][0], out[False][0]
    assert np.array_equal(A_rows.indptr, A_col.indptr) and np.array_equal(A_rows.indices, A_col.indices)
    assert np.array_equal(A_rows.data, A_col.data)
    assert np.array_equal(out[True][1], out[False][1])


def test_magnetostatics_rejects_partly_prescribed_nodes(gpu):
    """The reference's magnetostatic book is per node: a 3-DOF book that prescribes some
    components of a node and not others is refused on the host, before any launch."""
    rs = gpu
    xyz, tets = rs.kuhn_box(4, 4, 4)
    nn = xyz.shape[0]
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    fixed[3 * 31 + 2] = 1          # node 31 is at x = 0.25: free before
    asm = rs.Assembly(xyz, tets, 3, fixed)
    zeros = rs.DeviceArray.from_host(np.zeros(3 * nn))
    for rows in (True, False):
        asm.set_method(rows)
        with pytest.raises(rs.RSError, match="whole nodes"):
            asm.assemble_magnetostatics_device(zeros, zeros, rs.DeviceArray(asm.n_dof))


def test_magnetostatics_multilevel_coarse_jacobi(gpu):
    """The magnetostatic operator is semidefinite, and so can be its coarsest Galerkin operator
    (the exact coarse inverse refuses it at 96^3); the multilevel cycle over per-component node
    bins with coarse Jacobi on the coarsest level solves the system to Jacobi-CG's solution (1e-5)
    in a fraction of its iterations (no counterpart in the reference)."""
    rs = gpu
    c = 16
    xyz, tets = rs.kuhn_box(c, c, c)
    nn = xyz.shape[0]
    x = xyz[:, 0]
    fixed = np.repeat((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12), 3).astype(np.uint8)
    J = np.random.default_rng(11).standard_normal(3 * nn) * 1e6
    asm = rs.Assembly(xyz, tets, 3, fixed)
    n = asm.n_dof
    db = rs.DeviceArray(n)
    asm.assemble_magnetostatics_device(rs.DeviceArray.from_host(J), rs.DeviceArray.from_host(np.zeros(3 * nn)), db)
    dj = rs.DeviceArray.from_host(np.zeros(n))
    jac = rs.RMatrixSolver(rs.MatrixSolverConf(n_outer_iterations=100000, solver_cvg_value=1e-9,
                                               stop_rule=rs.STOP_RELRES)).solve_device(asm.matrix, db, dj,
                                                                                       rs.PRECOND_JACOBI, 1)
    assert jac.converged == 1
    x_jac = dj.download()
    L = rs.LearnedPreconditioner.from_bins(asm.matrix, asm, (c // 2, c // 2, c // 2), max_coarse=rs.MAX_COARSE,
                                           by_component=True)
    conf = rs.MatrixSolverConf(n_outer_iterations=5000, solver_cvg_value=1e-9, stop_rule=rs.STOP_RELRES,
                               precision=rs.PRECISION_MIXED)
    L.set_cycle(1, 1, rs.COARSE_JACOBI)
    dx = rs.DeviceArray.from_host(np.zeros(n))
    info = rs.RMatrixSolver(conf).solve_learned_device(asm.matrix, L, db, dx)
    assert info.converged == 1 and info.true_relres <= 1e-9
    x = dx.download()
    assert np.linalg.norm(x - x_jac) / np.linalg.norm(x_jac) <= 1e-5
    print(f"magnetostatics {c}^3: multilevel {info.iterations} its, Jacobi-CG {jac.iterations} its", flush=True)
    assert info.iterations * 3 < jac.iterations
