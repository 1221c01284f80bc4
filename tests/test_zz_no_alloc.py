"""The allocation-free FP64 CG (rs_solver_solve_device_noalloc) against rs_solver_solve_device.

Both run the same kernels in the same order on the same SpMV shape, so the solutions, iteration
counts and stopping data must be bit-identical.  The calls come in a row on one GPU, each
smaller than the one before, so every part of the shared workspace that the previous call left
behind is overwritten or not read."""
import numpy as np
import pytest

from helpers import heat_problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run(rs, method, M, b, precond, width, conf):
    S = rs.RMatrixSolver(conf)
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray.from_host(np.zeros_like(b))
    info = getattr(S, method)(M, db, dx, precond, width)
    return dx.download(), info


def _same(rs, M, b, precond, width, conf, name, forced_plan=None):
    if forced_plan is None:
        x_ref, i_ref = _run(rs, "solve_device", M, b, precond, width, conf)
        x, info = _run(rs, "solve_device_noalloc", M, b, precond, width, conf)
    else:
        # a matrix never tuned: the allocation-free call runs vector CSR; the other entry point
        # is held to the same shape
        x, info = _run(rs, "solve_device_noalloc", M, b, precond, width, conf)
        M.force_plan(rs.PRECISION_FP64, forced_plan)
        x_ref, i_ref = _run(rs, "solve_device", M, b, precond, width, conf)
    _say(f"{name}: n={b.size} its {info.iterations} (with allocation {i_ref.iterations}), "
         f"relres {info.true_relres:.2e}")
    assert info.converged == 1 and i_ref.converged == 1
    assert info.iterations == i_ref.iterations and info.reason == i_ref.reason
    assert info.error == i_ref.error and info.equation_scale == i_ref.equation_scale
    assert info.true_relres == i_ref.true_relres
    assert np.array_equal(x, x_ref)


def _heat(rs, cells):
    xyz, tets, fixed, t_node, k, q = heat_problem(cells, 2.0, rs=rs)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    b = asm.assemble_heat(k, q, t_node)
    return asm, b


def test_noalloc_matches_allocating_solve_on_calls_in_a_row(gpu):
    rs = gpu
    conf = rs.MatrixSolverConf(n_outer_iterations=20000, solver_cvg_value=1e-10)
    asm, b = _heat(rs, 20)
    _same(rs, asm.matrix, b, rs.PRECOND_JACOBI, 1, conf, "heat 20^3 Jacobi")
    small, bs = _heat(rs, 10)                       # smaller, never tuned
    _same(rs, small.matrix, bs, rs.PRECOND_JACOBI, 1, conf, "heat 10^3 Jacobi, untuned", forced_plan=2)
    xyz, tets = rs.kuhn_box(12, 3, 3, 4.0, 1.0, 1.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    tris = rs.plane_faces(xyz, tets, 0, 4.0)
    cant = rs.Assembly(xyz, tets, 3, fixed)
    bc = cant.assemble_stress(rs.field_lognormal(3, ne, 0.0, 1.0), np.full(ne, 0.3), np.ones(ne), [0.0, 0.0, 0.0],
                              np.zeros(3 * nn), tris, np.tile([0.0, 0.0, -1.0], (tris.shape[0], 1)))
    _same(rs, cant.matrix, bc, rs.PRECOND_BLOCK_JACOBI, 3, conf, "cantilever 12x3x3 block-Jacobi(3)")
    _same(rs, small.matrix, bs, rs.PRECOND_NONE, 1, conf, "heat 10^3 no preconditioner", forced_plan=2)


def test_noalloc_refuses_what_it_cannot_do(gpu):
    rs = gpu
    n = 3_400_000                                   # 4n + n doubles: above the 128 MiB workspace
    M = rs.SparseMatrix(n, np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), np.full(n, 2.0))
    db, dx = rs.DeviceArray.from_host(np.ones(n)), rs.DeviceArray(n)
    S = rs.RMatrixSolver(rs.MatrixSolverConf())
    with pytest.raises(rs.RSError) as e:
        S.solve_device_noalloc(M, db, dx, rs.PRECOND_JACOBI, 1)
    assert e.value.code == rs.RS_ERR_INVALID_ARG and "134217728" in str(e.value)
    asm, b = _heat(rs, 4)
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray(b.size)
    for conf in (rs.MatrixSolverConf(solver_type=rs.SOLVER_GMRES), rs.MatrixSolverConf(precision=rs.PRECISION_MIXED)):
        with pytest.raises(rs.RSError) as e:
            rs.RMatrixSolver(conf).solve_device_noalloc(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
        assert e.value.code == rs.RS_ERR_UNSUPPORTED
    with pytest.raises(rs.RSError) as e:
        S.solve_device_noalloc(asm.matrix, db, dx, 7, 1)
    assert e.value.code == rs.RS_ERR_INVALID_PRECOND
