"""GPU tests of the learned multilevel preconditioner (it has no counterpart in the reference).

What pins what:
  * level operators against the scipy Galerkin chain P^T A P (an independent formulation) and
    the smoother bounds against numpy Gershgorin sums;
  * one V-cycle against oracle/learned_oracle.py's cycle (itself pinned by closed forms in
    tests/test_learned_oracle.py) on the same level operators;
  * the inference kernel and the training kernel against the oracle's network evaluated with
    the same BF16 rounding points; that oracle's backward pass is pinned by central finite
    differences, and each comparison checks that a planted mistake in the oracle (a transposed
    layer matrix, a wrong smoothing count) lands far outside the tolerance;
  * the flexible-CG solve against an FP64 Jacobi-CG solution, with the FP64 true residual.
Tolerances: level operators 1e-12 * max|A| (FP64 sums in another order); V-cycle 1e-4 relative
in the 2-norm (FP32 vectors and values against FP64); network outputs and gradients 2e-2
relative (BF16 activations: 2^-8 relative per rounding, and a rounding that flips between
FP32 and FP64 accumulation moves one BF16 ulp).
"""
import numpy as np
import pytest

import learned_oracle as lo
from helpers import heat_problem

pytestmark = pytest.mark.gpu

NET_RTOL = 2e-2
CYC_RTOL = 1e-4


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300)


def _heat(rs, n, sigma, bins=None):
    xyz, tets, fixed, t_node, k, q = heat_problem(n, sigma, rs=rs)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    b = asm.assemble_heat(k, q, t_node)
    if bins is None:
        return asm, b, xyz
    agg, nagg = rs.aggregates_by_bins(xyz, asm.book(), 1, bins)
    return asm, b, agg, nagg


def _theta(seed, head=0.3):
    """Random hidden layers and a non-zero output head (a zero head would leave every hidden
    gradient zero and make the output independent of the layers)."""
    rng = np.random.default_rng(seed)
    g = lo.unpack(lo.init_params(seed))
    g["wo"] = rng.standard_normal(lo.HID) * head / np.sqrt(lo.HID)
    g["bo"] = np.array([0.05])
    g["bin"] = 0.1 * rng.standard_normal(lo.HID)
    for l in range(lo.LAYERS):
        g["b"][l] = 0.05 * rng.standard_normal(lo.HID)
    return lo.pack(g)


def _transposed(theta):
    g = lo.unpack(theta)
    g["Ws"] = [w.T.copy() for w in g["Ws"]]
    return lo.pack(g)


def _levels_from_gpu(L):
    return [L.level_matrix(l) for l in range(L.n_levels)]


def test_aggregates_by_bins(gpu):
    rs = gpu
    asm, _, agg, nagg = _heat(rs, 8, 0.0, (2, 2, 2))
    assert nagg == 8 and agg.size == asm.n_dof
    assert set(np.unique(agg)) == set(range(8))
    counts = np.bincount(agg, minlength=8)
    assert counts.min() > 0 and counts.sum() == asm.n_dof


@pytest.mark.parametrize("bins,max_coarse", [((8, 8, 8), 64), ((6, 6, 6), 512)])
def test_hierarchy_operators_match_galerkin_chain(gpu, bins, max_coarse):
    rs = gpu
    _say(f"hierarchy bins={bins} max_coarse={max_coarse}")
    asm, _, xyz = _heat(rs, 16, 1.5)
    maps = rs.hierarchy_by_bins(xyz, asm.book(), 1, bins, max_coarse=max_coarse)
    L = rs.LearnedPreconditioner(asm.matrix, maps, params=lo.init_params(0))
    L.setup()
    A = asm.matrix.to_scipy()
    ref, _ = lo.hierarchy(A, maps)
    assert L.n_levels == len(maps) + 1
    scale = abs(A.data).max()
    for l in range(L.n_levels):
        G = L.level_matrix(l)
        assert G.shape == ref[l].shape
        assert abs(G - ref[l]).max() <= 1e-12 * scale, l
    np.testing.assert_allclose(L.level_lambda(), [lo.gershgorin(M) for M in ref], rtol=1e-12)
    np.testing.assert_allclose(L.coarse_matrix(), ref[-1].toarray(), atol=1e-12 * scale)
    # deterministic: a second hierarchy on the same matrix is bit-identical
    L2 = rs.LearnedPreconditioner(asm.matrix, maps, params=lo.init_params(0))
    L2.setup()
    for l in range(L.n_levels):
        assert np.array_equal(L2.level_matrix(l).data, L.level_matrix(l).data)


def _stress_assembly(rs, nx, ny, nz):
    xyz, tets = rs.kuhn_box(nx, ny, nz, 4.0, 1.0, 1.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    E = rs.field_lognormal(3, ne, 0.0, 1.0)
    tris = rs.plane_faces(xyz, tets, 0, 4.0)
    trac = np.tile([0.0, 0.0, -1.0], (tris.shape[0], 1))
    asm = rs.Assembly(xyz, tets, 3, fixed)
    b = asm.assemble_stress(E, np.full(ne, 0.3), np.ones(ne), [0.0, 0.0, 0.0], np.zeros(3 * nn), tris, trac)
    return asm, b, xyz


@pytest.mark.parametrize("case", ["heat16-8^3-64", "heat20-10x7x5-200", "stress12x4x4-6^3-100",
                                  "stress12x4x4-components-6x2x2-100"])
def test_hierarchy_built_on_gpu(gpu, case):
    """rs_learned_create_by_bins: the level maps built on the GPU are the oracle's nested bins,
    and every level operator and bound is bit-identical to those of the hierarchy built from the
    host maps; a learned solve with either converges alike."""
    rs = gpu
    _say(f"GPU hierarchy {case}")
    if case.startswith("heat16"):
        (asm, b, xyz), dpn, bins, mc = _heat(rs, 16, 1.5), 1, (8, 8, 8), 64
    elif case.startswith("heat20"):
        (asm, b, xyz), dpn, bins, mc = _heat(rs, 20, 2.0), 1, (10, 7, 5), 200
    elif "components" in case:
        (asm, b, xyz), dpn, bins, mc = _stress_assembly(rs, 12, 4, 4), 3, (6, 2, 2), 100
    else:
        (asm, b, xyz), dpn, bins, mc = _stress_assembly(rs, 12, 4, 4), 3, (6, 6, 6), 100
    comp = "components" in case
    ref = lo.bins_hierarchy(xyz, asm.book(), dpn, bins, mc, by_component=comp)
    Lg = rs.LearnedPreconditioner.from_bins(asm.matrix, asm, bins, max_coarse=mc, params=lo.init_params(1),
                                            by_component=comp)
    maps = rs.hierarchy_by_bins(xyz, asm.book(), dpn, bins, max_coarse=mc, by_component=comp)
    Lh = rs.LearnedPreconditioner(asm.matrix, maps, params=lo.init_params(1))
    assert Lg.n_levels == len(ref) + 1
    for l, r in enumerate(ref):
        assert np.array_equal(Lg.level_map(l), r), l
    assert np.array_equal(Lg.level_rows, Lh.level_rows) and np.array_equal(Lg.level_nnz, Lh.level_nnz)
    Lg.setup()
    Lh.setup()
    for l in range(Lg.n_levels):
        G, H = Lg.level_matrix(l), Lh.level_matrix(l)
        assert np.array_equal(G.indptr, H.indptr) and np.array_equal(G.indices, H.indices), l
        assert np.array_equal(G.data, H.data), l
    assert np.array_equal(Lg.level_lambda(), Lh.level_lambda())
    if dpn == 1:
        # each preconditioner autotunes its own coarse levels' FP32 SpMV shapes (timing-based),
        # so the two solves may sum coarse rows in different orders: same bar, not the same bits
        xg, ig = _learned_solve(rs, asm, b, Lg)
        xh, ih = _learned_solve(rs, asm, b, Lh)
        assert ig.converged == 1 and ih.converged == 1 and abs(ig.iterations - ih.iterations) <= 1
        assert _rel(xg, xh) <= 1e-6


def test_hierarchy_on_gpu_invalid_arguments(gpu):
    rs = gpu
    asm, _, _ = _heat(rs, 6, 0.0)
    with pytest.raises(rs.RSError):
        rs.LearnedPreconditioner.from_bins(asm.matrix, asm, (4, 4, 4), max_coarse=600)
    with pytest.raises(rs.RSError):   # one level allowed, but 4^3 bins leave more than 8 nodes
        rs.LearnedPreconditioner.from_bins(asm.matrix, asm, (4, 4, 4), max_coarse=8, max_levels=1)
    other, _, _ = _heat(rs, 5, 0.0)   # a book that does not number this matrix's rows
    with pytest.raises(rs.RSError):
        rs.LearnedPreconditioner.from_bins(asm.matrix, other, (4, 4, 4), max_coarse=64)


def test_rigid_hierarchy_invalid_arguments(gpu):
    """RS_AGG_RIGID needs 3 DOFs per node, and refuses aggregates of more than 256 rows."""
    rs = gpu
    asm, _, _ = _heat(rs, 6, 0.0)
    with pytest.raises(rs.RSError):   # 1 DOF per node
        rs.LearnedPreconditioner.from_bins(asm.matrix, asm, (3, 3, 3), max_coarse=64, rigid=True)
    sasm, _, _ = _stress_assembly(rs, 16, 4, 4)
    with pytest.raises(rs.RSError):   # bins of 8 x 5 x 5 nodes: 600 rows per aggregate
        rs.LearnedPreconditioner.from_bins(sasm.matrix, sasm, (2, 1, 1), max_coarse=512, rigid=True)
    L = rs.LearnedPreconditioner.from_bins(sasm.matrix, sasm, (8, 2, 2), max_coarse=512, rigid=True)
    assert L.level_rows[1] % 6 == 0 and L.level_rows[-1] <= 512
    with pytest.raises(rs.RSError):   # not a rigid-body level: no basis rows
        rs.LearnedPreconditioner.from_bins(sasm.matrix, sasm, (8, 2, 2), max_coarse=512).level_basis(0)


def test_single_level_coarse_operator(gpu):
    rs = gpu
    asm, _, agg, nagg = _heat(rs, 12, 1.0, (3, 3, 4))
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L.setup()
    A = asm.matrix.to_scipy()
    np.testing.assert_allclose(L.coarse_matrix(), lo.galerkin(A, agg, nagg), atol=1e-12 * abs(A.data).max())


def test_restriction_to_coarsest(gpu):
    rs = gpu
    asm, b, xyz = _heat(rs, 12, 1.0)
    maps = rs.hierarchy_by_bins(xyz, asm.book(), 1, (6, 6, 6), max_coarse=30)
    L = rs.LearnedPreconditioner(asm.matrix, maps)
    total = np.arange(asm.n_dof)
    for m in maps:
        total = m[total]
    r = np.random.default_rng(4).standard_normal(asm.n_dof)
    ref = np.bincount(total, weights=r, minlength=L.n_coarse)
    np.testing.assert_allclose(L.restrict(rs.DeviceArray.from_host(r)), ref, rtol=1e-6, atol=1e-6 * abs(ref).max())


def _coarse_fn(mode):
    return lo.coarse_exact if mode == "exact" else lo.coarse_jacobi


@pytest.mark.parametrize("pre,post,mode,fp16", [(1, 1, "jacobi", False), (2, 2, "exact", False),
                                                (0, 1, "exact", False), (3, 2, "jacobi", False),
                                                (1, 1, "jacobi", True), (2, 1, "exact", True)])
def test_vcycle_matches_oracle(gpu, pre, post, mode, fp16):
    rs = gpu
    _say(f"vcycle pre={pre} post={post} coarse={mode} fp16-smoother={fp16}")
    asm, _, xyz = _heat(rs, 20, 2.0)
    maps = rs.hierarchy_by_bins(xyz, asm.book(), 1, (10, 10, 10), max_coarse=200)
    L = rs.LearnedPreconditioner(asm.matrix, maps)
    L.set_cycle(pre, post, rs.COARSE_EXACT if mode == "exact" else rs.COARSE_JACOBI)
    L.set_compressed(1 if fp16 else 0)
    r = np.random.default_rng(pre * 10 + post).standard_normal(asm.n_dof).astype(np.float32)
    z = L.apply(r)
    levels = _levels_from_gpu(L)
    _, Ps = lo.hierarchy(asm.matrix.to_scipy(), maps)
    lams = L.level_lambda()
    fp16_levels = tuple(range(len(levels) - 1)) if fp16 else ()
    ref = lo.vcycle(levels, Ps, r.astype(np.float64), pre, post, _coarse_fn(mode), lams, fp16_levels=fp16_levels)
    assert np.all(np.isfinite(z))
    assert _rel(z, ref) < CYC_RTOL, _rel(z, ref)
    # a planted mistake (one post-smoothing step more) is far outside the tolerance
    bad = lo.vcycle(levels, Ps, r.astype(np.float64), pre, post + 1, _coarse_fn(mode), lams, fp16_levels=fp16_levels)
    assert _rel(bad, ref) > 100 * CYC_RTOL


@pytest.mark.parametrize("pre,post,cpre,cpost,mode", [(1, 1, 3, 3, "jacobi"), (1, 1, 0, 2, "exact"),
                                                      (2, 1, 1, 4, "exact")])
def test_vcycle_coarse_degrees_match_oracle(gpu, pre, post, cpre, cpost, mode):
    """Smoothing steps on the coarse levels set apart from the finest (rs_learned_set_coarse_cycle)."""
    rs = gpu
    _say(f"vcycle pre={pre} post={post} coarse pre={cpre} post={cpost} coarse={mode}")
    asm, _, xyz = _heat(rs, 20, 2.0)
    maps = rs.hierarchy_by_bins(xyz, asm.book(), 1, (10, 10, 10), max_coarse=200)
    L = rs.LearnedPreconditioner(asm.matrix, maps)
    L.set_cycle(pre, post, rs.COARSE_EXACT if mode == "exact" else rs.COARSE_JACOBI)
    L.set_coarse_cycle(cpre, cpost)
    L.set_compressed(0)
    r = np.random.default_rng(cpre * 10 + cpost).standard_normal(asm.n_dof).astype(np.float32)
    z = L.apply(r)
    levels = _levels_from_gpu(L)
    assert len(levels) >= 3
    _, Ps = lo.hierarchy(asm.matrix.to_scipy(), maps)
    lams = L.level_lambda()
    ref = lo.vcycle(levels, Ps, r.astype(np.float64), pre, post, _coarse_fn(mode), lams, coarse_steps=(cpre, cpost))
    assert np.all(np.isfinite(z))
    assert _rel(z, ref) < CYC_RTOL, _rel(z, ref)
    # the same cycle with the finest level's steps everywhere is far outside the tolerance
    bad = lo.vcycle(levels, Ps, r.astype(np.float64), pre, post, _coarse_fn(mode), lams)
    assert _rel(bad, ref) > 100 * CYC_RTOL


def test_inference_matches_oracle(gpu):
    rs = gpu
    _say("inference 8x8x8 aggregates")
    asm, _, agg, nagg = _heat(rs, 24, 2.0, (8, 8, 8))
    assert nagg == 512
    theta = _theta(7)
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=theta)
    L.setup()
    Ac = L.coarse_matrix()
    rng = np.random.default_rng(8)
    rc = (Ac @ rng.standard_normal(nagg)).astype(np.float32)
    e_gpu = L.apply_coarse(rc).astype(np.float64)
    u = rc / np.diag(Ac)
    sigma = np.sqrt(np.mean(u * u))
    e_ref, cache = lo.forward(theta, Ac, rc.astype(np.float64), bf16_points=True)
    assert np.all(np.isfinite(e_gpu))
    assert _rel((e_gpu - u) / sigma, cache["o"]) < NET_RTOL
    assert _rel(e_gpu, e_ref) < NET_RTOL
    _, bad = lo.forward(_transposed(theta), Ac, rc.astype(np.float64), bf16_points=True)
    assert _rel(bad["o"], cache["o"]) > 5 * NET_RTOL
    # the untrained network (zero head) is coarse Jacobi
    L.set_params(lo.init_params(3))
    np.testing.assert_allclose(L.apply_coarse(rc), u, rtol=1e-5)


def _samples(Ac, B, seed):
    """Coarse residuals of smooth errors e = A_c^-1 w, with s = r.A_c^-1.r = e.r."""
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((B, Ac.shape[0]))
    E = np.linalg.solve(Ac, W.T).T
    R = E @ Ac.T
    return R, np.einsum("bi,bi->b", E, R)


def _check_gradient(theta, Ac, R32, S32, loss, grad):
    B = R32.shape[0]
    g_ref = np.zeros_like(theta)
    l_ref = np.zeros(B)
    for i in range(B):
        g, l_ref[i] = lo.backward(theta, Ac, R32[i].astype(np.float64), float(S32[i]), bf16_points=True)
        g_ref += g / B
    np.testing.assert_allclose(loss, l_ref, rtol=1e-3, atol=1e-4)
    assert _rel(grad, g_ref) < NET_RTOL, _rel(grad, g_ref)
    gg, gr = lo.unpack(grad), lo.unpack(g_ref)
    for key in ("Win", "bin", "wo", "bo"):
        assert _rel(gg[key], gr[key]) < 2 * NET_RTOL, key
    for l in range(lo.LAYERS):
        for key in ("Ws", "Wn", "b"):
            assert _rel(gg[key][l], gr[key][l]) < 2 * NET_RTOL, (key, l)
    g_bad = sum(lo.backward(_transposed(theta), Ac, R32[i].astype(np.float64), float(S32[i]),
                            bf16_points=True)[0] for i in range(B)) / B
    assert _rel(g_bad, g_ref) > 10 * NET_RTOL


@pytest.mark.parametrize("bins", [(4, 4, 4), (3, 3, 4)])
def test_trainer_gradient_dense_matches_oracle(gpu, bins):
    rs = gpu
    _say(f"trainer gradient (dense pattern) bins={bins}")
    asm, _, agg, nagg = _heat(rs, 12, 1.5, bins)
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L.setup()
    Ac = L.coarse_matrix().astype(np.float32).astype(np.float64)
    B = 6
    R, S = _samples(Ac, B, 11)
    R32, S32 = R.astype(np.float32), S.astype(np.float32)
    theta = _theta(12)
    T = rs.Trainer(theta, max_batch=16, max_nodes=64, max_nnz=64 * 64)
    loss, grad = T.gradient(np.broadcast_to(Ac, (B, nagg, nagg)), R32, S32)
    _check_gradient(theta, Ac, R32, S32, loss, grad)


def test_trainer_gradient_sparse_512_matches_oracle(gpu):
    rs = gpu
    _say("trainer gradient (coarse pattern) 512 nodes")
    asm, _, agg, nagg = _heat(rs, 24, 2.0, (8, 8, 8))
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L.setup()
    C = L.level_matrix(1)
    C.data = C.data.astype(np.float32).astype(np.float64)
    Ac = C.toarray()
    B = 4
    R, S = _samples(Ac, B, 13)
    R32, S32 = R.astype(np.float32), S.astype(np.float32)
    theta = _theta(14)
    T = rs.Trainer(theta, max_batch=8, max_nodes=512, max_nnz=C.nnz)
    cvals = np.broadcast_to(C.data.astype(np.float32), (B, C.nnz))
    loss, grad = T.gradient_csr(C.indptr, C.indices, cvals, R32, S32)
    _check_gradient(theta, Ac, R32, S32, loss, grad)


def test_trainer_adam_reduces_loss(gpu):
    rs = gpu
    asm, _, agg, nagg = _heat(rs, 12, 1.5, (4, 4, 4))
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L.setup()
    Ac = L.coarse_matrix().astype(np.float32).astype(np.float64)
    B = 32
    R, S = _samples(Ac, B, 21)
    theta0 = lo.init_params(5)
    T = rs.Trainer(theta0, max_batch=B, max_nodes=64, max_nnz=64 * 64)
    A_b = np.broadcast_to(Ac, (B, nagg, nagg))
    l0, _ = T.gradient(A_b, R.astype(np.float32), S.astype(np.float32))
    # untrained: the loss is the A-norm error fraction left by coarse Jacobi
    e0 = R / np.diag(Ac)
    frac = np.array([1 + (e @ Ac @ e - 2 * e @ r) / s for e, r, s in zip(e0, R, S)])
    np.testing.assert_allclose(l0, frac, rtol=1e-3)
    losses = [T.step(A_b, R.astype(np.float32), S.astype(np.float32), 2e-3) for _ in range(60)]
    _say(f"adam: loss {np.mean(l0):.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.5 * np.mean(l0)
    theta = T.params()
    assert theta.shape == theta0.shape and not np.array_equal(theta, theta0)
    L.set_params(theta)
    e_gpu = L.apply_coarse(R[0].astype(np.float32))
    e_ref, _ = lo.forward(theta, Ac, R[0].astype(np.float32).astype(np.float64), bf16_points=True)
    assert _rel(e_gpu, e_ref) < NET_RTOL


def _fp64_solution(rs, M, b):
    conf = rs.MatrixSolverConf(n_outer_iterations=20000, solver_cvg_value=1e-13, stop_rule=rs.STOP_RELRES)
    S = rs.RMatrixSolver(conf)
    x = S.solve(M, b, None, rs.PRECOND_JACOBI, 1)
    return x, S.info.iterations


def _learned_solve(rs, asm, b, L):
    conf = rs.MatrixSolverConf(n_outer_iterations=5000, solver_cvg_value=1e-8, stop_rule=rs.STOP_RELRES,
                               precision=rs.PRECISION_MIXED)
    S = rs.RMatrixSolver(conf)
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray.from_host(np.zeros_like(b))
    info = S.solve_learned_device(asm.matrix, L, db, dx)
    return dx.download(), info


@pytest.fixture(scope="module")
def heat24(gpu):
    rs = gpu
    asm, b, xyz = _heat(rs, 24, 2.0)
    maps = rs.hierarchy_by_bins(xyz, asm.book(), 1, (12, 12, 12), max_coarse=512)
    x_star, jac_its = _fp64_solution(rs, asm.matrix, b)
    return asm, b, maps, x_star, jac_its


@pytest.mark.parametrize("mode", ["network_untrained", "network_random", "jacobi", "exact"])
def test_learned_fcg_solve(gpu, heat24, mode):
    rs = gpu
    asm, b, maps, x_star, jac_its = heat24
    theta = _theta(2, head=1.0) if mode == "network_random" else lo.init_params(1)
    L = rs.LearnedPreconditioner(asm.matrix, maps, params=theta)
    coarse = {"exact": rs.COARSE_EXACT, "jacobi": rs.COARSE_JACOBI}.get(mode, rs.COARSE_NETWORK)
    L.set_cycle(1, 1, coarse)
    x, info = _learned_solve(rs, asm, b, L)
    _say(f"learned FCG {mode}: its={info.iterations} fallback={info.fallback} relres={info.true_relres:.2e} "
         f"(Jacobi-CG {jac_its})")
    assert info.converged == 1, info.as_dict()
    assert info.true_relres <= 1e-8
    A = asm.matrix.to_scipy()
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1.5e-8
    assert np.linalg.norm(x - x_star) / np.linalg.norm(x_star) < 1e-6
    if mode != "network_random":
        assert info.fallback == 0
        assert info.iterations < 0.4 * jac_its


def test_exact_coarse_not_worse_than_coarse_jacobi(gpu, heat24):
    rs = gpu
    asm, b, maps, _, _ = heat24
    its = {}
    for mode in (rs.COARSE_JACOBI, rs.COARSE_EXACT):
        L = rs.LearnedPreconditioner(asm.matrix, maps)
        L.set_cycle(1, 1, mode)
        its[mode] = _learned_solve(rs, asm, b, L)[1].iterations
    assert its[rs.COARSE_EXACT] <= its[rs.COARSE_JACOBI]


@pytest.mark.parametrize("cells,bins", [(24, (12, 12, 12)), (16, (8, 8, 8))], ids=["216_padded", "512"])
def test_exact_coarse_inverse_on_device(gpu, cells, bins):
    """The coarsest operator's inverse computed on the GPU (blocked sweep in FP64, stored in
    FP32) against numpy's dense solve, on a size that is padded to a multiple of 64 (216) and on
    512.  Coarse Jacobi, a planted stand-in for a wrong inverse, lands far outside the bar."""
    rs = gpu
    asm, _, xyz = _heat(rs, cells, 1.5)
    maps = rs.hierarchy_by_bins(xyz, asm.book(), 1, bins, max_coarse=512)
    L = rs.LearnedPreconditioner(asm.matrix, maps)
    L.set_cycle(1, 1, rs.COARSE_EXACT)
    L.setup()
    Ac = L.coarse_matrix()
    n = L.n_coarse
    _say(f"exact coarse inverse on the device: n_coarse={n}")
    rng = np.random.default_rng(11)
    rc = rng.standard_normal(n).astype(np.float32)
    e = L.apply_coarse(rc).astype(np.float64)
    e_ref = np.linalg.solve(Ac, rc.astype(np.float64))
    assert _rel(e, e_ref) < 1e-5
    assert np.linalg.norm(Ac @ e - rc) / np.linalg.norm(rc) < 1e-4
    L.set_cycle(1, 1, rs.COARSE_JACOBI)
    assert _rel(L.apply_coarse(rc), e_ref) > 1e-2


def test_records_are_the_cycle_coarsest_rhs(gpu, heat24):
    rs = gpu
    asm, b, maps, _, _ = heat24
    L = rs.LearnedPreconditioner(asm.matrix, maps, params=lo.init_params(1))
    L.set_cycle(1, 1, rs.COARSE_NETWORK)
    L.record(8)
    x, info = _learned_solve(rs, asm, b, L)
    rec = L.records(min(info.iterations, 8))
    assert rec.shape == (min(info.iterations, 8), L.n_coarse) and np.all(np.isfinite(rec))
    # iteration 0: the cycle is applied to the scaled right-hand side (zero start)
    r = (b * info.equation_scale).astype(np.float32).astype(np.float64)
    levels = _levels_from_gpu(L)
    _, Ps = lo.hierarchy(asm.matrix.to_scipy(), maps)
    lams = L.level_lambda()
    for l in range(L.n_levels - 1):
        A = levels[l]
        z = lo.cheb_smooth(A, 1.0 / A.diagonal(), lo.cheb_coefs(lams[l]), r, None, 1)
        r = Ps[l].T @ (r - A @ z)
    assert _rel(rec[0], r) < 1e-3


def test_learned_fallback_on_non_finite_network(gpu, heat24):
    """A network whose output is not finite: the flexible CG stops on the non-finite residual
    and Jacobi-CG finishes from the last good iterate."""
    rs = gpu
    asm, b, maps, _, _ = heat24
    g = lo.unpack(lo.init_params(1))
    g["bo"] = np.array([np.nan])
    L = rs.LearnedPreconditioner(asm.matrix, maps, params=lo.pack(g))
    x, info = _learned_solve(rs, asm, b, L)
    assert info.fallback == 1
    assert info.converged == 1 and info.true_relres <= 1e-8
    assert np.all(np.isfinite(x))


def test_learned_invalid_arguments(gpu):
    rs = gpu
    asm, b, agg, nagg = _heat(rs, 6, 0.0, (2, 2, 2))
    with pytest.raises(rs.RSError):
        rs.LearnedPreconditioner(asm.matrix, agg, 600)        # beyond 512 coarsest nodes
    bad = agg.copy()
    bad[0] = nagg
    with pytest.raises(rs.RSError):
        rs.LearnedPreconditioner(asm.matrix, bad, nagg)       # aggregate id out of range
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg)
    with pytest.raises(rs.RSError):
        L.set_cycle(1, 0, rs.COARSE_NETWORK)                  # no post-smoothing
    with pytest.raises(rs.RSError):
        L.set_cycle(1, 1, 7)                                   # unknown coarse mode
    S = rs.RMatrixSolver(rs.MatrixSolverConf())
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray(b.size)
    with pytest.raises(rs.RSError):
        S.solve_learned_device(asm.matrix, L, db, dx)         # network parameters not set
    other = rs.SparseMatrix.from_scipy(asm.matrix.to_scipy())
    L.set_params(lo.init_params(0))
    with pytest.raises(rs.RSError):
        L.setup(other)                                         # built for another matrix
    with pytest.raises(rs.RSError):
        rs.Trainer(np.zeros(10, np.float32))                   # wrong parameter count
    T = rs.Trainer(lo.init_params(0), max_batch=2, max_nodes=8, max_nnz=64)
    cptr = np.array([0, 2, 3], np.uint32)
    ccol = np.array([0, 1, 1], np.uint32)                      # (0,1) without (1,0)
    with pytest.raises(rs.RSError):
        T.gradient_csr(cptr, ccol, np.ones((1, 3), np.float32), np.ones((1, 2), np.float32),
                       np.ones(1, np.float32))
