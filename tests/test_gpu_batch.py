"""GPU parity tests of the batched solver (rs_batch_*): many uneven systems, each running its
own implicit-Euler steps of Jacobi-CG, all sharing every launch.

The oracle is the reference's time loop restated per system (oracle.transient_heat: assemble
with the previous temperatures, Jacobi-CG from x = 0, write back), run to 1e-13.  Bars: FP64
batch solutions within 1e-8 of the oracle and per-step iterations within 2 of the oracle's
CG at the same tolerance (the batch stops on the FP64 true residual, the oracle on its
recursive one); MIXED batch solutions within 1e-6 with every step's true relative residual
below the tolerance.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import random_spd

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def tr(gpu):
    import __graft_entry__
    return __graft_entry__.load_module("transient")


def _conf(rs, precision, cvg, n_outer=20000):
    return rs.MatrixSolverConf(n_outer_iterations=n_outer, solver_cvg_value=cvg, stop_rule=rs.STOP_RELRES,
                               precision=precision)


def _oracle_system(rs, orc, tr, meshes, sizes, i, seed, sigma, dt, steps, cvg):
    m = meshes[sizes[i]]
    k, q = tr.fields(rs, m, seed, i, sigma)
    book, n_dof = orc.node_book(m.xyz.shape[0], m.tets, 1, m.fixed)
    return orc.transient_heat(m.xyz, m.tets, k, q, np.ones(m.tets.shape[0]), m.t0, book, n_dof, dt, steps, cvg=cvg)


SIZES = [3, 5, 2, 4, 3, 6]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_transient_matches_oracle(gpu, orc, tr, precision):
    rs = gpu
    sigma, dt, steps, seed = 1.5, 1e-2, 4, 9
    prec = rs.PRECISION_FP64 if precision == "fp64" else rs.PRECISION_MIXED
    cvg = 1e-10 if precision == "fp64" else 1e-8
    _say(f"batch transient {precision}: sizes {SIZES}, {steps} steps")
    B, meshes, _, _ = tr.build_batch(rs, SIZES, sigma=sigma, dt=dt, seed=seed)
    info = B.solve(_conf(rs, prec, cvg), steps)
    its, rr, reasons = B.stats()
    assert info.systems_converged == len(SIZES), reasons
    assert (reasons == 1).all()
    assert (rr <= cvg).all() and info.max_relres <= cvg
    assert info.total_iterations == its.sum()
    for i in range(len(SIZES)):
        xs, _ = _oracle_system(rs, orc, tr, meshes, SIZES, i, seed, sigma, dt, steps, 1e-13)
        x = B.solution(i)
        err = np.linalg.norm(x - xs[-1]) / np.linalg.norm(xs[-1])
        assert err <= (1e-8 if precision == "fp64" else 1e-6), (i, err)
        if precision == "fp64":
            _, oits = _oracle_system(rs, orc, tr, meshes, SIZES, i, seed, sigma, dt, steps, cvg)
            assert np.abs(its[i] - np.array(oits)).max() <= 2, (i, its[i], oits)


def test_batch_steady_random_spd(gpu):
    """Ragged steady systems (no mass matrix, one step) of 0..300 rows, including sizes that
    are exact multiples of the 64-row slice, against a direct solve."""
    rs = gpu
    ns = [0, 1, 5, 63, 64, 65, 128, 200, 300, 17]
    mats, bs, handles = [], [], []
    for j, n in enumerate(ns):
        if n == 0:
            mats.append(None)
            bs.append(None)
            continue
        A = random_spd(n, 0.05, 100 + j)
        A.sort_indices()
        mats.append(A)
        bs.append(np.random.default_rng(j).standard_normal(n))
    B = rs.Batch([n for n in ns], [0 if A is None else A.nnz for A in mats])
    for i, (A, b) in enumerate(zip(mats, bs)):
        if A is None:
            continue
        M = rs.SparseMatrix.from_scipy(A)
        handles.append(M)
        B.set_matrix(i, M)
        db = rs.DeviceArray.from_host(b)
        handles.append(db)
        B.set_rhs(i, db)
    for prec, cvg, tol in ((rs.PRECISION_FP64, 1e-12, 1e-9), (rs.PRECISION_MIXED, 1e-9, 1e-6)):
        info = B.solve(_conf(rs, prec, cvg), 1)
        its, rr, reasons = B.stats()
        assert (reasons == 1).all(), reasons
        for i, (A, b) in enumerate(zip(mats, bs)):
            if A is None:
                assert its[i, 0] == 0
                continue
            x = B.solution(i)
            xr = np.linalg.solve(A.toarray(), b)
            assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= tol, (prec, i)
            assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= cvg
        assert info.rounds >= its.max()


def test_batch_zero_rhs_and_initial_solution(gpu):
    """A zero right-hand side converges in zero iterations; a nonzero initial solution enters
    step 0 through the mass matrix (b_0 = c + M x_init)."""
    rs = gpu
    A = random_spd(100, 0.05, 1)
    A.sort_indices()
    Mm = A.copy()                                 # same pattern as A, values 2 I
    rows = np.repeat(np.arange(100), np.diff(A.indptr))
    Mm.data = np.where(rows == A.indices, 2.0, 0.0)
    B = rs.Batch([100, 100], [A.nnz, A.nnz])
    dA, dM = rs.SparseMatrix.from_scipy(A), rs.SparseMatrix.from_scipy(Mm)
    x_init = np.linspace(0.0, 1.0, 100)
    c = np.zeros(100)
    dc, dx0 = rs.DeviceArray.from_host(c), rs.DeviceArray.from_host(x_init)
    zero = rs.DeviceArray.from_host(np.zeros(100))
    for i in range(2):
        B.set_matrix(i, dA)
        B.set_mass(i, dM)
    B.set_rhs(0, dc)             # zero everything: x stays 0
    B.set_rhs(1, dc, dx0)
    B.solve(_conf(rs, rs.PRECISION_FP64, 1e-12), 3)
    its, rr, reasons = B.stats()
    assert (its[0] == 0).all() and (reasons == 1).all()
    assert np.abs(B.solution(0)).max() == 0.0
    Ad = A.toarray()
    x = x_init.copy()
    for _ in range(3):
        x = np.linalg.solve(Ad, 2.0 * x)
    xb = B.solution(1)
    assert np.linalg.norm(xb - x) / np.linalg.norm(x) <= 1e-9
    del zero


def test_batch_warm_start_shapes_and_determinism(gpu, tr):
    rs = gpu
    sizes = [4, 7, 3, 6, 5, 8, 2, 7]
    B, _, _, _ = tr.build_batch(rs, sizes, sigma=1.5, dt=1e-2, seed=21)
    conf = _conf(rs, rs.PRECISION_MIXED, 1e-8)
    info = B.solve(conf, 5)
    its0, _, _ = B.stats()
    xs0 = [B.solution(i) for i in range(len(sizes))]
    # bit-identical on a second run
    B.solve(conf, 5)
    its1, _, _ = B.stats()
    assert np.array_equal(its0, its1)
    for i in range(len(sizes)):
        assert np.array_equal(xs0[i], B.solution(i))
    # warm starts: the same solutions to the tolerance, fewer iterations in total
    winfo = B.solve(conf, 5, warm_start=True)
    assert winfo.total_iterations < info.total_iterations
    for i in range(len(sizes)):
        xw = B.solution(i)
        assert np.linalg.norm(xw - xs0[i]) / np.linalg.norm(xs0[i]) <= 1e-6
    # every SpMV shape gives the same answers to the tolerance
    shape, ms = B.shape(rs.PRECISION_MIXED)
    assert 0 <= shape < rs.BATCH_SHAPES and (ms > 0).all()
    for sid in range(rs.BATCH_SHAPES):
        B.force_shape(rs.PRECISION_MIXED, sid)
        B.solve(conf, 5)
        its_s, rr_s, reasons = B.stats()
        assert (reasons == 1).all() and (rr_s <= 1e-8).all(), sid
        assert np.abs(its_s.astype(int) - its0.astype(int)).max() <= 3, sid
        for i in range(len(sizes)):
            x = B.solution(i)
            assert np.linalg.norm(x - xs0[i]) / np.linalg.norm(xs0[i]) <= 1e-6, (sid, i)
    B.force_shape(rs.PRECISION_MIXED, -1)


def test_batch_invalid_arguments(gpu):
    rs = gpu
    A = random_spd(70, 0.05, 2)
    A.sort_indices()
    other = random_spd(70, 0.05, 3)
    other.sort_indices()
    dA = rs.SparseMatrix.from_scipy(A)
    B = rs.Batch([70, 70], [A.nnz, A.nnz])
    B.set_matrix(0, dA)
    with pytest.raises(rs.RSError):
        B.solve(_conf(rs, rs.PRECISION_FP64, 1e-10), 1)          # system 1: no matrix
    B.set_matrix(1, dA)
    with pytest.raises(rs.RSError):
        B.solve(_conf(rs, rs.PRECISION_FP64, 1e-10), 1)          # no right-hand sides
    db = rs.DeviceArray.from_host(np.ones(70))
    B.set_rhs(0, db)
    B.set_rhs(1, db)
    with pytest.raises(rs.RSError):
        B.solve(_conf(rs, rs.PRECISION_FP64, 1e-10), 0)          # zero steps
    with pytest.raises(rs.RSError):
        B.solve(rs.MatrixSolverConf(solver_cvg_value=1e-10), 1)  # reference stopping rule
    with pytest.raises(rs.RSError):
        B.set_mass(0, rs.SparseMatrix.from_scipy(other))         # wrong size / pattern
    B.set_mass(0, dA)
    with pytest.raises(rs.RSError):
        B.solve(_conf(rs, rs.PRECISION_FP64, 1e-10), 1)          # mass for one system only
    with pytest.raises(rs.RSError):
        rs.Batch([], [])
    B.set_mass(1, dA)
    info = B.solve(_conf(rs, rs.PRECISION_FP64, 1e-10), 2)
    assert info.systems_converged == 2


def test_batch_column_offsets_match_full_columns(gpu, tr, monkeypatch):
    """int16 column offsets (the default on mesh matrices) and full uint32 columns read the same
    entries in the same order: bit-identical solves."""
    rs = gpu
    sizes = [5, 3, 6, 4]
    out = {}
    for full in (False, True):
        if full:
            monkeypatch.setenv("RS_BATCH_FULL_COLUMNS", "1")
        B, _, _, _ = tr.build_batch(rs, sizes, sigma=1.5, dt=1e-2, seed=33)
        for prec in (rs.PRECISION_FP64, rs.PRECISION_MIXED):
            B.solve(_conf(rs, prec, 1e-9), 3)
            out[(full, prec)] = (B.stats()[0].copy(), [B.solution(i) for i in range(len(sizes))])
        assert B.layout()["column_bytes"] == (4 if full else 2)
    for prec in (rs.PRECISION_FP64, rs.PRECISION_MIXED):
        a, b = out[(False, prec)], out[(True, prec)]
        assert np.array_equal(a[0], b[0])
        for x, y in zip(a[1], b[1]):
            assert np.array_equal(x, y)


def test_batch_replacement_ratio(gpu, tr):
    """Fewer FP64 replacements with a smaller ratio; the steps still end below the tolerance."""
    rs = gpu
    B, _, _, _ = tr.build_batch(rs, [6, 4, 7], sigma=1.5, dt=1e-2, seed=41)
    conf = _conf(rs, rs.PRECISION_MIXED, 1e-8)
    reps = []
    for ratio in (0.1, 1e-2, 1e-3):
        B.set_replacement(ratio)
        info = B.solve(conf, 3)
        its, rr, reasons = B.stats()
        assert (reasons == 1).all() and (rr <= 1e-8).all(), ratio
        reps.append(info.replacements)
    assert reps[0] > reps[1] >= reps[2] >= 9      # at least one per step and system
    B.set_replacement(0.01)
    with pytest.raises(rs.RSError):
        B.set_replacement(1.0)

