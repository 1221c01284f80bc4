"""GPU parity of the configured workloads against the oracle, at reduced sizes and at full size.

Steady heat (the 128^3 headline's problem: conductivity lognormal(0, 2) from seed 2, source
U(0, 1) from seed 1, T = 0 on x = 0 and T = 1 on x = 1) and the elasticity cantilever (box
4 x 1 x 1, E lognormal(0, 1) from seed 3, nu = 0.3, clamped at x = 0, traction (0, 0, -1) on
x = 4) are assembled on the GPU and by the oracle, then:
  * the CSR pattern is bit-exact and the values and right-hand side agree within
    1e-12 * max|.| (FP64 sums in another order);
  * FP64 Jacobi-CG (heat) or block-Jacobi(3) CG (cantilever) to relres 1e-8 takes the oracle's
    number of iterations within +-1 (the oracle restates src/rmatrixsolver.cpp:208-355 and
    src/rmatrixpreconditioner.cpp:96-190 in the reference's order);
  * mixed precision (FP32 SpMV and preconditioner, FP64 residual replacement) and the learned
    preconditioner in flexible CG reach the FP64 true relres 1e-8 with a solution error
    ||x - x*|| / ||x*|| <= 1e-6, x* the oracle's FP64 solution converged to relres 1e-13.
"""
import numpy as np
import pytest

from helpers import heat_problem

pytestmark = pytest.mark.gpu

ERR_BAR = 1e-6
RELRES = 1e-8


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _same_system(A_gpu, b_gpu, A_ref, b_ref):
    assert A_gpu.shape == A_ref.shape
    assert np.array_equal(A_gpu.indptr, A_ref.indptr) and np.array_equal(A_gpu.indices, A_ref.indices)
    assert np.abs(A_gpu.data - A_ref.data).max() <= 1e-12 * np.abs(A_ref.data).max()
    assert np.abs(b_gpu - b_ref).max() <= 1e-12 * np.abs(b_ref).max()


def _gpu_solve(rs, M, b, precision, precond, width, cvg=RELRES):
    conf = rs.MatrixSolverConf(n_outer_iterations=200000, solver_cvg_value=cvg, stop_rule=rs.STOP_RELRES,
                               precision=precision)
    S = rs.RMatrixSolver(conf)
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray.from_host(np.zeros_like(b))
    info = S.solve_device(M, db, dx, precond, width)
    return dx.download(), info


def _learned_solve(rs, M, b, L, precision=None):
    conf = rs.MatrixSolverConf(n_outer_iterations=5000, solver_cvg_value=RELRES, stop_rule=rs.STOP_RELRES,
                               precision=rs.PRECISION_MIXED if precision is None else precision)
    S = rs.RMatrixSolver(conf)
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray.from_host(np.zeros_like(b))
    info = S.solve_learned_device(M, L, db, dx)
    return dx.download(), info


def _oracle_solution(orc, A, b, precond, width, history=False):
    """x* (relres 1e-13); with history, also the iteration at which the oracle's own stopping
    rule (src/riterationinfo.cpp:90-109) would have stopped at relres 1e-8 on the same run."""
    x, info, hist = orc.solve(A, b, precond=precond, block_size=width, n_outer=200000, cvg=1e-13,
                              stop_rule=orc.STOP_RELRES, history=True)
    # the recursive residual drifts from the true one on the ill-conditioned cantilever: x* is
    # held to true relres 1e-10, two orders below the 1e-8 it is a reference for
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
    if not history:
        return x
    its_1e8 = next(i + 1 for i, (e, t) in enumerate(hist) if orc.converged_reason(e, t, RELRES) != 0)
    return x, its_1e8


def _check_converged(A, b, x, info):
    assert info.converged == 1, info.as_dict()
    assert info.true_relres <= RELRES
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1.5 * RELRES


class HeatCase:
    def __init__(self, rs, orc, cells, history=False):
        _say(f"heat {cells}^3, lognormal(0, 2): GPU and oracle assembly")
        self.cells = cells
        xyz, tets, fixed, t_node, k, q = heat_problem(cells, 2.0, seed_k=2, seed_q=1, rs=rs)
        self.xyz = xyz
        self.asm = rs.Assembly(xyz, tets, 1, fixed)
        self.b = self.asm.assemble_heat(k, q, t_node)
        book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
        assert np.array_equal(book, self.asm.book()) and n_dof == self.asm.n_dof
        self.A_ref, self.b_ref = orc.heat_assemble(xyz, tets, k, q, t_node, book, n_dof)
        self.A = self.asm.matrix.to_scipy()
        _say(f"heat {cells}^3: oracle FP64 Jacobi-CG to relres 1e-13")
        if history:
            self.x_star, self.oracle_its = _oracle_solution(orc, self.A_ref, self.b_ref, orc.PRECOND_JACOBI, 1, True)
        else:
            self.x_star = _oracle_solution(orc, self.A_ref, self.b_ref, orc.PRECOND_JACOBI, 1)


@pytest.fixture(scope="module", params=[24, 32])
def heat(request, gpu, orc):
    return HeatCase(gpu, orc, request.param)


def test_heat_assembly_matches_oracle(heat):
    _same_system(heat.A, heat.b, heat.A_ref, heat.b_ref)


def test_heat_fp64_jacobi_cg_iterations_match_oracle(gpu, orc, heat):
    rs = gpu
    x, info = _gpu_solve(rs, heat.asm.matrix, heat.b, rs.PRECISION_FP64, rs.PRECOND_JACOBI, 1)
    _, oinfo = orc.solve(heat.A_ref, heat.b_ref, precond=orc.PRECOND_JACOBI, n_outer=200000, cvg=RELRES,
                         stop_rule=orc.STOP_RELRES)
    _say(f"heat {heat.cells}^3 FP64 Jacobi-CG: {info.iterations} its (oracle {oinfo['iterations']})")
    assert abs(info.iterations - oinfo["iterations"]) <= 1
    _check_converged(heat.A, heat.b, x, info)
    assert _rel(x, heat.x_star) <= ERR_BAR


def test_heat_mixed_cg_against_oracle_solution(gpu, heat):
    rs = gpu
    x, info = _gpu_solve(rs, heat.asm.matrix, heat.b, rs.PRECISION_MIXED, rs.PRECOND_JACOBI, 1)
    err = _rel(x, heat.x_star)
    _say(f"heat {heat.cells}^3 mixed Jacobi-CG: {info.iterations} its, error vs oracle {err:.2e}")
    _check_converged(heat.A, heat.b, x, info)
    assert err <= ERR_BAR


@pytest.mark.parametrize("variant", ["default", "fp16_smoother", "fp16_smoother_vecrows_fp32", "fp64_krylov"])
def test_heat_learned_fcg_against_oracle_solution(gpu, heat, variant):
    """The trained network in the flexible CG.  `fp16_smoother` forces the fine level's FP16
    smoother copy below its default size threshold; with `vecrows_fp32` the FP32 SpMV plan is
    also forced to a non-SELL shape, so the smoother and the SpMV run different grids (the
    flexible CG's r.z and z.q partial sums come from the smoother's grid); `fp64_krylov` runs
    the flexible CG's vectors and product in FP64 around the FP32 cycle (PRECISION_FP64)."""
    rs = gpu
    params, _ = rs.load_weights()
    assert params is not None, "the committed weights are missing"
    maps = rs.hierarchy_by_bins(heat.xyz, heat.asm.book(), 1, (heat.cells // 2,) * 3, max_coarse=rs.MAX_COARSE)
    M = heat.asm.matrix
    L = rs.LearnedPreconditioner(M, maps, params=params)
    L.set_cycle(1, 1, rs.COARSE_NETWORK)
    if variant != "default":
        L.set_compressed(1)
    if variant == "fp16_smoother_vecrows_fp32":
        M.force_plan(rs.PRECISION_MIXED, 1)
    try:
        x, info = _learned_solve(rs, M, heat.b, L, rs.PRECISION_FP64 if variant == "fp64_krylov" else None)
        if variant == "fp16_smoother_vecrows_fp32":
            assert M.spmv_plan(rs.PRECISION_MIXED)[0] == 1
    finally:
        M.force_plan(rs.PRECISION_MIXED, -1)
    err = _rel(x, heat.x_star)
    _say(f"heat {heat.cells}^3 learned FCG ({variant}): {info.iterations} its, fallback {info.fallback}, "
         f"error vs oracle {err:.2e}")
    _check_converged(heat.A, heat.b, x, info)
    assert info.fallback == 0
    assert err <= ERR_BAR


class CantileverCase:
    def __init__(self, rs, orc, cells):
        nx, ny, nz = cells
        _say(f"cantilever {nx}x{ny}x{nz}: GPU and oracle assembly")
        self.cells = cells
        xyz, tets = rs.kuhn_box(nx, ny, nz, 4.0, 1.0, 1.0)
        nn, ne = xyz.shape[0], tets.shape[0]
        fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
        E = rs.field_lognormal(3, ne, 0.0, 1.0)
        nu, rho = np.full(ne, 0.3), np.ones(ne)
        tris = rs.plane_faces(xyz, tets, 0, 4.0)
        trac = np.tile([0.0, 0.0, -1.0], (tris.shape[0], 1))
        self.asm = rs.Assembly(xyz, tets, 3, fixed)
        self.b = self.asm.assemble_stress(E, nu, rho, [0.0, 0.0, 0.0], np.zeros(3 * nn), tris, trac)
        book, n_dof = orc.node_book(nn, tets, 3, fixed)
        assert np.array_equal(book, self.asm.book()) and n_dof == self.asm.n_dof
        self.A_ref, self.b_ref = orc.stress_assemble(xyz, tets, E, nu, rho, [0.0, 0.0, 0.0], np.zeros(3 * nn), tris,
                                                     trac, book, n_dof)
        self.A = self.asm.matrix.to_scipy()
        _say(f"cantilever {nx}x{ny}x{nz}: oracle FP64 block-Jacobi(3) CG to relres 1e-13")
        self.x_star = _oracle_solution(orc, self.A_ref, self.b_ref, orc.PRECOND_BLOCK_JACOBI, 3)


@pytest.fixture(scope="module", params=[(16, 4, 4), (32, 8, 8)], ids=["16x4x4", "32x8x8"])
def cantilever(request, gpu, orc):
    return CantileverCase(gpu, orc, request.param)


def test_cantilever_assembly_matches_oracle(cantilever):
    _same_system(cantilever.A, cantilever.b, cantilever.A_ref, cantilever.b_ref)


def test_cantilever_fp64_blockjacobi_cg_iterations_match_oracle(gpu, orc, cantilever):
    rs = gpu
    c = cantilever
    x, info = _gpu_solve(rs, c.asm.matrix, c.b, rs.PRECISION_FP64, rs.PRECOND_BLOCK_JACOBI, 3)
    _, oinfo = orc.solve(c.A_ref, c.b_ref, precond=orc.PRECOND_BLOCK_JACOBI, block_size=3, n_outer=200000,
                         cvg=RELRES, stop_rule=orc.STOP_RELRES)
    _say(f"cantilever {c.cells} FP64 block-Jacobi(3) CG: {info.iterations} its (oracle {oinfo['iterations']})")
    assert abs(info.iterations - oinfo["iterations"]) <= 1
    _check_converged(c.A, c.b, x, info)
    assert _rel(x, c.x_star) <= ERR_BAR


def test_cantilever_mixed_blockjacobi_cg_against_oracle_solution(gpu, cantilever):
    rs = gpu
    c = cantilever
    x, info = _gpu_solve(rs, c.asm.matrix, c.b, rs.PRECISION_MIXED, rs.PRECOND_BLOCK_JACOBI, 3)
    err = _rel(x, c.x_star)
    _say(f"cantilever {c.cells} mixed block-Jacobi(3) CG: {info.iterations} its, {info.inner_iterations} "
         f"replacements, error vs oracle {err:.2e}")
    _check_converged(c.A, c.b, x, info)
    assert err <= ERR_BAR


def test_cantilever_multilevel_fcg_against_oracle_solution(gpu, cantilever):
    """Elasticity through the multilevel cycle: per-component nested bins (unknown-based
    aggregation, RS_AGG_COMPONENTS), one Chebyshev step before and after, the exact coarsest
    inverse, flexible CG in mixed precision.  The level maps are the oracle's; the solution is
    within 1e-6 of the oracle's FP64 solution; the iterations are at most twice those of the
    oracle's FP64 flexible CG with the same cycle (learned_oracle.vcycle, FP64; the GPU's Krylov
    vectors, SpMV and cycle run in FP32, which the cantilever's cancellation amplifies: 125
    against 84 at 16x4x4), and far below block-Jacobi(3) CG's."""
    import learned_oracle as lo
    rs = gpu
    c = cantilever
    nx, ny, nz = c.cells
    bins = (max(1, nx // 2), max(1, ny // 2), max(1, nz // 2))
    L = rs.LearnedPreconditioner.from_bins(c.asm.matrix, c.asm, bins, max_coarse=rs.MAX_COARSE, by_component=True)
    L.set_cycle(1, 1, rs.COARSE_EXACT)
    xyz = rs.kuhn_box(nx, ny, nz, 4.0, 1.0, 1.0)[0]
    maps = lo.bins_hierarchy(xyz, c.asm.book(), 3, bins, rs.MAX_COARSE, by_component=True)
    assert L.n_levels == len(maps) + 1
    for l, m in enumerate(maps):
        assert np.array_equal(L.level_map(l), m), l
    x, info = _learned_solve(rs, c.asm.matrix, c.b, L)
    err = _rel(x, c.x_star)
    x64, info64 = _learned_solve(rs, c.asm.matrix, c.b, L, rs.PRECISION_FP64)
    err64 = _rel(x64, c.x_star)
    levels, Ps = lo.hierarchy(c.A_ref, maps)
    lams = [lo.gershgorin(A) for A in levels[:-1]]
    cinv = np.linalg.inv(levels[-1].toarray())
    fp16 = tuple(l for l in range(len(levels) - 1) if levels[l].shape[0] >= 65536)
    _, oits = lo.fcg(c.A_ref, c.b_ref, lambda r: lo.vcycle(levels, Ps, r, 1, 1, lambda A, rc: cinv @ rc, lams, 0,
                                                            fp16), tol=RELRES)
    _, bj = _gpu_solve(rs, c.asm.matrix, c.b, rs.PRECISION_FP64, rs.PRECOND_BLOCK_JACOBI, 3)
    _say(f"cantilever {c.cells} multilevel FCG (per-component bins, exact coarsest): {info.iterations} its, "
         f"FP64 Krylov {info64.iterations} its (oracle FP64 cycle {oits}, block-Jacobi(3) CG {bj.iterations}), "
         f"levels {list(map(int, L.level_rows))}, error vs oracle {err:.2e} / {err64:.2e}")
    _check_converged(c.A, c.b, x, info)
    assert info.fallback == 0
    assert err <= ERR_BAR
    assert info.iterations <= 2 * oits
    assert 2 * info.iterations < bj.iterations
    # FP64 Krylov vectors around the FP32 cycle: the oracle's iterations (the cycle's FP32
    # rounding only perturbs the preconditioner, which the flexible CG absorbs)
    _check_converged(c.A, c.b, x64, info64)
    assert info64.fallback == 0
    assert err64 <= ERR_BAR
    assert abs(info64.iterations - oits) <= 0.1 * oits + 2


def test_cantilever_rigid_multilevel_fcg_against_oracle(gpu, cantilever):
    """Elasticity through the rigid-body hierarchy (RS_AGG_RIGID: node aggregates whose coarse
    spaces are their six rigid-body modes, built on the GPU) against the oracle's restatement
    (learned_oracle.rigid_hierarchy): level maps equal; basis rows and level operators within
    1e-12; flexible CG with FP64 Krylov vectors takes the oracle's FP64 iterations (within 10% + 2)
    and mixed precision at most twice as many, both within 1e-6 of the oracle's FP64 solution."""
    import learned_oracle as lo
    rs = gpu
    c = cantilever
    nx, ny, nz = c.cells
    bins = (max(1, nx // 2), max(1, ny // 2), max(1, nz // 2))
    _say(f"cantilever {c.cells}: rigid-body hierarchy on the GPU")
    L = rs.LearnedPreconditioner.from_bins(c.asm.matrix, c.asm, bins, max_coarse=rs.MAX_COARSE, rigid=True)
    L.set_cycle(1, 1, rs.COARSE_EXACT)
    xyz = rs.kuhn_box(nx, ny, nz, 4.0, 1.0, 1.0)[0]
    book = c.asm.book()
    levels, Ps, deads = lo.rigid_hierarchy(c.A_ref, xyz, book, bins, rs.MAX_COARSE)
    maps = lo.bins_hierarchy(xyz, book, 3, bins, rs.MAX_COARSE // 6)
    assert L.n_levels == len(levels)
    assert [int(v) for v in L.level_rows] == [Lv.shape[0] for Lv in levels]
    L.setup()
    for l, P in enumerate(Ps):
        rowmap = maps[l] if l == 0 else np.repeat(maps[l], 6)
        assert np.array_equal(L.level_map(l), rowmap), l
        q = L.level_basis(l)
        Pd = P.tocsr()
        ref = np.zeros_like(q)
        for a in range(6):
            ref[:, a] = np.asarray(Pd[np.arange(q.shape[0]), 6 * rowmap.astype(np.int64) + a]).ravel()
        assert np.abs(q - ref).max() <= 1e-12, l
        G, R = L.level_matrix(l + 1), levels[l + 1]
        assert np.abs(G - R).max() <= 1e-12 * np.abs(R).max(), l
    lams = [lo.gershgorin(Lv) for Lv in levels[:-1]]
    cinv = np.linalg.inv(levels[-1].toarray())
    fp16 = tuple(l for l in range(len(levels) - 1) if levels[l].shape[0] >= 65536)
    _, oits = lo.fcg(c.A_ref, c.b_ref, lambda r: lo.vcycle(levels, Ps, r, 1, 1, lambda A, rc: cinv @ rc, lams, 0,
                                                            fp16), tol=RELRES)
    x64, i64 = _learned_solve(rs, c.asm.matrix, c.b, L, rs.PRECISION_FP64)
    xmp, imp = _learned_solve(rs, c.asm.matrix, c.b, L)
    e64, emp = _rel(x64, c.x_star), _rel(xmp, c.x_star)
    _say(f"cantilever {c.cells} rigid multilevel FCG: FP64 Krylov {i64.iterations} its, mixed {imp.iterations} its "
         f"(oracle FP64 cycle {oits}), levels {list(map(int, L.level_rows))}, error vs oracle {e64:.2e} / {emp:.2e}")
    for x, info, err in ((x64, i64, e64), (xmp, imp, emp)):
        _check_converged(c.A, c.b, x, info)
        assert info.fallback == 0
        assert err <= ERR_BAR
    assert abs(i64.iterations - oits) <= 0.1 * oits + 2
    assert imp.iterations <= 2 * oits


def test_trained_network_beats_coarse_jacobi_on_held_out_problem(gpu, orc):
    """The committed weights (trained on 64^3 problems with lognormal(0, 2.5) conductivity, seeds
    10000-11999) on a problem none of them saw (seed 20002): fewer flexible-CG iterations than
    the same cycle with coarse Jacobi on the coarsest level (the untrained network computes
    exactly that), and both solutions within 1e-6 of the oracle's."""
    rs = gpu
    cells = 64
    xyz, tets, fixed, t_node, k, q = heat_problem(cells, 2.5, seed_k=20002, seed_q=1, rs=rs)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    b = asm.assemble_heat(k, q, t_node)
    A = asm.matrix.to_scipy()
    _say("held-out 64^3 sigma 2.5: oracle FP64 Jacobi-CG to relres 1e-13")
    x_star = _oracle_solution(orc, A, b, orc.PRECOND_JACOBI, 1)
    params, _ = rs.load_weights()
    maps = rs.hierarchy_by_bins(xyz, asm.book(), 1, (cells // 2,) * 3, max_coarse=rs.MAX_COARSE)
    its = {}
    for name, coarse in (("network", rs.COARSE_NETWORK), ("coarse_jacobi", rs.COARSE_JACOBI)):
        L = rs.LearnedPreconditioner(asm.matrix, maps, params=params)
        L.set_cycle(1, 1, coarse)
        x, info = _learned_solve(rs, asm.matrix, b, L)
        _check_converged(A, b, x, info)
        err = _rel(x, x_star)
        _say(f"held-out 64^3 {name}: {info.iterations} its, error vs oracle {err:.2e}")
        assert err <= ERR_BAR
        its[name] = info.iterations
    assert its["network"] < its["coarse_jacobi"], its


@pytest.fixture(scope="module")
def heat128(gpu, orc):
    return HeatCase(gpu, orc, 128, history=True)


def test_heat128_assembly_and_fp64_cg_match_oracle(gpu, heat128):
    """Full size (2,113,407 unknowns): the GPU system and the FP64 Jacobi-CG iteration count
    against the oracle's."""
    rs = gpu
    h = heat128
    _same_system(h.A, h.b, h.A_ref, h.b_ref)
    x, info = _gpu_solve(rs, h.asm.matrix, h.b, rs.PRECISION_FP64, rs.PRECOND_JACOBI, 1)
    _say(f"heat 128^3 FP64 Jacobi-CG: {info.iterations} its (oracle {h.oracle_its})")
    assert abs(info.iterations - h.oracle_its) <= 1
    _check_converged(h.A, h.b, x, info)
    assert _rel(x, h.x_star) <= ERR_BAR


def test_heat128_mixed_and_learned_against_oracle_solution(gpu, heat128):
    rs = gpu
    h = heat128
    x, info = _gpu_solve(rs, h.asm.matrix, h.b, rs.PRECISION_MIXED, rs.PRECOND_JACOBI, 1)
    err = _rel(x, h.x_star)
    _say(f"heat 128^3 mixed Jacobi-CG: {info.iterations} its, error vs oracle {err:.2e}")
    _check_converged(h.A, h.b, x, info)
    assert err <= ERR_BAR
    params, _ = rs.load_weights()
    maps = rs.hierarchy_by_bins(h.xyz, h.asm.book(), 1, (64, 64, 64), max_coarse=rs.MAX_COARSE)
    L = rs.LearnedPreconditioner(h.asm.matrix, maps, params=params)
    L.set_cycle(1, 1, rs.COARSE_NETWORK)
    x, info = _learned_solve(rs, h.asm.matrix, h.b, L)
    err = _rel(x, h.x_star)
    _say(f"heat 128^3 learned FCG: {info.iterations} its, error vs oracle {err:.2e}")
    _check_converged(h.A, h.b, x, info)
    assert info.fallback == 0
    assert err <= ERR_BAR
