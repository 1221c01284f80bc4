"""GPU tests of the learned two-level preconditioner (it has no counterpart in the reference).

What pins what:
  * the coarse operator against scipy's P^T A P (an independent formulation);
  * the inference kernel and the training kernel against oracle/learned_oracle.py evaluated
    with the same BF16 rounding points; that oracle's backward pass is pinned by central finite
    differences (tests/test_learned_oracle.py), and each comparison here also checks that a
    planted mistake in the oracle (a transposed layer matrix) is far outside the tolerance;
  * the flexible-CG solve against an FP64 Jacobi-CG solution, with the FP64 true residual.
Tolerances: coarse operator 1e-10 * max|A| (64-bit fixed point, 2^-44 * max|A| per term);
network outputs and gradients 2e-2 relative in the 2-norm (BF16 activations: 2^-8 relative
per rounding, a flipped rounding between FP32 and FP64 accumulation moves one ulp).
"""
import numpy as np
import pytest

import learned_oracle as lo
from helpers import heat_problem

pytestmark = pytest.mark.gpu

NET_RTOL = 2e-2


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300)


def _heat(rs, n, sigma, bins):
    xyz, tets, fixed, t_node, k, q = heat_problem(n, sigma, rs=rs)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    b = asm.assemble_heat(k, q, t_node)
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


def test_aggregates_by_bins(gpu):
    rs = gpu
    asm, _, agg, nagg = _heat(rs, 8, 0.0, (2, 2, 2))
    assert nagg == 8 and agg.size == asm.n_dof
    assert set(np.unique(agg)) == set(range(8))
    counts = np.bincount(agg, minlength=8)
    assert counts.min() > 0 and counts.sum() == asm.n_dof


@pytest.mark.parametrize("bins", [(4, 4, 4), (3, 3, 4)])
def test_coarse_operator_matches_galerkin(gpu, bins):
    rs = gpu
    _say(f"galerkin bins={bins}")
    asm, _, agg, nagg = _heat(rs, 12, 1.0, bins)
    A = asm.matrix.to_scipy()
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L.setup()
    Ac = L.coarse_matrix()
    ref = lo.galerkin(A, agg, nagg)
    assert Ac.shape == (nagg, nagg)
    assert abs(Ac - ref).max() <= 1e-10 * abs(A.data).max()
    assert np.array_equal(Ac != 0, ref != 0) or abs(ref[(Ac != 0) != (ref != 0)]).max() < 1e-9
    # deterministic: a second set-up (fresh object) is bit-identical
    L2 = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L2.setup()
    assert np.array_equal(L2.coarse_matrix(), Ac)


def test_restriction_and_pair(gpu):
    rs = gpu
    asm, b, agg, nagg = _heat(rs, 10, 1.0, (3, 3, 3))
    A = asm.matrix.to_scipy()
    n = asm.n_dof
    rng = np.random.default_rng(4)
    x = rng.standard_normal(n)
    xs = rng.standard_normal(n)
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L.setup()
    db, dx, dxs = rs.DeviceArray.from_host(b), rs.DeviceArray.from_host(x), rs.DeviceArray.from_host(xs)
    rc, s = L.pair(asm.matrix, db, dx, dxs)
    r = b - A @ x
    rc_ref = np.bincount(agg, weights=r, minlength=nagg)
    np.testing.assert_allclose(rc, rc_ref, rtol=1e-6, atol=1e-6 * abs(rc_ref).max())
    assert abs(s - (xs - x) @ r) <= 1e-10 * abs(xs - x) @ abs(r)
    dr = rs.DeviceArray.from_host(r)
    np.testing.assert_allclose(L.restrict(dr), rc_ref, rtol=1e-6, atol=1e-6 * abs(rc_ref).max())


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
    e_true = rng.standard_normal(nagg)
    rc = (Ac @ e_true).astype(np.float32)
    e_gpu = L.apply_coarse(rc).astype(np.float64)
    d = np.diag(Ac)
    u = rc / d
    sigma = np.sqrt(np.mean(u * u))
    o_gpu = (e_gpu - u) / sigma
    e_ref, cache = lo.forward(theta, Ac, rc.astype(np.float64), bf16_points=True)
    o_ref = cache["o"]
    assert np.all(np.isfinite(e_gpu))
    assert _rel(o_gpu, o_ref) < NET_RTOL, _rel(o_gpu, o_ref)
    assert _rel(e_gpu, e_ref) < NET_RTOL
    # the tolerance separates a planted mistake
    _, bad = lo.forward(_transposed(theta), Ac, rc.astype(np.float64), bf16_points=True)
    assert _rel(bad["o"], o_ref) > 5 * NET_RTOL
    # the untrained network (zero head) is coarse Jacobi
    L.set_params(lo.init_params(3))
    np.testing.assert_allclose(L.apply_coarse(rc), u, rtol=1e-5)


def _samples(Ac, B, seed):
    """(residual, error) pairs of a smooth error: e_k = A_c^-1 w, rc = A_c e_k."""
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((B, Ac.shape[0]))
    E = np.linalg.solve(Ac, W.T).T
    R = E @ Ac.T
    S = np.einsum("bi,bi->b", E, R)
    return R, S


@pytest.mark.parametrize("bins", [(4, 4, 4), (3, 3, 4)])
def test_trainer_gradient_matches_oracle(gpu, bins):
    rs = gpu
    _say(f"trainer gradient bins={bins}")
    asm, _, agg, nagg = _heat(rs, 12, 1.5, bins)
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L.setup()
    Ac32 = L.coarse_matrix().astype(np.float32)
    Ac = Ac32.astype(np.float64)
    B = 6
    R, S = _samples(Ac, B, 11)
    R32, S32 = R.astype(np.float32), S.astype(np.float32)
    theta = _theta(12)
    T = rs.Trainer(theta, max_batch=16)
    loss, grad = T.gradient(np.broadcast_to(Ac32, (B, nagg, nagg)), R32, S32)
    g_ref = np.zeros_like(theta)
    l_ref = np.zeros(B)
    for i in range(B):
        g, l_ref[i] = lo.backward(theta, Ac, R32[i].astype(np.float64), float(S32[i]), bf16_points=True)
        g_ref += g / B
    np.testing.assert_allclose(loss, l_ref, rtol=1e-3, atol=1e-4)
    assert _rel(grad, g_ref) < NET_RTOL, _rel(grad, g_ref)
    # every parameter group, not just the dominant one
    gg, gr = lo.unpack(grad), lo.unpack(g_ref)
    for key in ("Win", "bin", "wo", "bo"):
        assert _rel(gg[key], gr[key]) < 2 * NET_RTOL, key
    for l in range(lo.LAYERS):
        for key in ("Ws", "Wn", "b"):
            assert _rel(gg[key][l], gr[key][l]) < 2 * NET_RTOL, (key, l)
    # planted mistake in the oracle: far outside the tolerance
    g_bad = sum(lo.backward(_transposed(theta), Ac, R32[i].astype(np.float64), float(S32[i]),
                            bf16_points=True)[0] for i in range(B)) / B
    assert _rel(g_bad, g_ref) > 10 * NET_RTOL


def test_trainer_adam_reduces_loss(gpu):
    rs = gpu
    asm, _, agg, nagg = _heat(rs, 12, 1.5, (4, 4, 4))
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.init_params(0))
    L.setup()
    Ac32 = L.coarse_matrix().astype(np.float32)
    Ac = Ac32.astype(np.float64)
    B = 32
    R, S = _samples(Ac, B, 21)
    theta0 = lo.init_params(5)
    T = rs.Trainer(theta0, max_batch=B)
    A_b = np.broadcast_to(Ac32, (B, nagg, nagg))
    l0, _ = T.gradient(A_b, R.astype(np.float32), S.astype(np.float32))
    # untrained: the loss is the A-norm error fraction left by coarse Jacobi
    e0 = R / np.diag(Ac)
    frac = np.array([1 + (e @ Ac @ e - 2 * e @ r) / s for e, r, s in zip(e0, R, S)])
    np.testing.assert_allclose(l0, frac, rtol=1e-3)
    losses = [T.step(A_b, R.astype(np.float32), S.astype(np.float32), 2e-3) for _ in range(60)]
    _say(f"adam: loss {np.mean(l0):.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.8 * np.mean(l0)
    theta = T.params()
    assert theta.shape == theta0.shape and not np.array_equal(theta, theta0)
    # the trained parameters, loaded into the inference kernel, match the oracle's forward
    L.set_params(theta)
    e_gpu = L.apply_coarse(R[0].astype(np.float32))
    e_ref, _ = lo.forward(theta, Ac, R[0].astype(np.float32).astype(np.float64), bf16_points=True)
    assert _rel(e_gpu, e_ref) < NET_RTOL


def _fp64_solution(rs, M, b):
    conf = rs.MatrixSolverConf(n_outer_iterations=20000, solver_cvg_value=1e-13, stop_rule=rs.STOP_RELRES)
    return rs.RMatrixSolver(conf).solve(M, b, None, rs.PRECOND_JACOBI, 1)


@pytest.mark.parametrize("which", ["untrained", "random_head"])
def test_learned_fcg_solve(gpu, which):
    rs = gpu
    _say(f"learned FCG solve {which}")
    asm, b, agg, nagg = _heat(rs, 24, 2.0, (6, 6, 6))
    theta = lo.init_params(1) if which == "untrained" else _theta(2, head=1.0)
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=theta)
    conf = rs.MatrixSolverConf(n_outer_iterations=5000, solver_cvg_value=1e-8, stop_rule=rs.STOP_RELRES,
                               precision=rs.PRECISION_MIXED)
    S = rs.RMatrixSolver(conf)
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray.from_host(np.zeros_like(b))
    info = S.solve_learned_device(asm.matrix, L, db, dx)
    x = dx.download()
    _say(f"  its={info.iterations} fallback={info.fallback} relres={info.true_relres:.2e}")
    assert info.converged == 1, info.as_dict()
    assert info.true_relres <= 1e-8
    A = asm.matrix.to_scipy()
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1.5e-8
    x_star = _fp64_solution(rs, asm.matrix, b)
    assert np.linalg.norm(x - x_star) / np.linalg.norm(x_star) < 1e-6
    assert info.precond_applications >= 1
    if which == "untrained":
        assert info.fallback == 0


def test_learned_fallback_on_non_finite_network(gpu):
    """A network whose output overflows: the flexible CG stops on the non-finite residual and
    Jacobi-CG finishes from the last good iterate."""
    rs = gpu
    asm, b, agg, nagg = _heat(rs, 16, 1.0, (4, 4, 4))
    g = lo.unpack(lo.init_params(1))
    g["bo"] = np.array([np.nan])
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg, params=lo.pack(g))
    conf = rs.MatrixSolverConf(n_outer_iterations=5000, solver_cvg_value=1e-8, stop_rule=rs.STOP_RELRES,
                               precision=rs.PRECISION_MIXED)
    S = rs.RMatrixSolver(conf)
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray.from_host(np.zeros_like(b))
    info = S.solve_learned_device(asm.matrix, L, db, dx)
    assert info.fallback == 1
    assert info.converged == 1 and info.true_relres <= 1e-8
    x = dx.download()
    assert np.all(np.isfinite(x))


def test_learned_invalid_arguments(gpu):
    rs = gpu
    asm, b, agg, nagg = _heat(rs, 6, 0.0, (2, 2, 2))
    with pytest.raises(rs.RSError):
        rs.LearnedPreconditioner(asm.matrix, agg, 600)        # beyond 512 aggregates
    bad = agg.copy()
    bad[0] = nagg
    with pytest.raises(rs.RSError):
        rs.LearnedPreconditioner(asm.matrix, bad, nagg)       # aggregate id out of range
    L = rs.LearnedPreconditioner(asm.matrix, agg, nagg)
    S = rs.RMatrixSolver(rs.MatrixSolverConf())
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray(b.size)
    with pytest.raises(rs.RSError):
        S.solve_learned_device(asm.matrix, L, db, dx)         # parameters not set
    with pytest.raises(rs.RSError):
        rs.Trainer(np.zeros(10, np.float32))                   # wrong parameter count
