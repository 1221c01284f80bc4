"""GPU parity tests of REigenValueSolver (rs_eigen_solve) against the oracle's restatement,
both started from the C library's rand() stream after the same srand(): eigenvalues to 1e-7
relative and vectors to 1e-5; plus the FEM Laplacian's first eigenvalue against pi^2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _heat_pair(rs, n, sigma=1.0, seed=5):
    """K: steady heat stiffness (lognormal conductivity), M: consistent mass (dt = 0), T fixed on
    x = 0 and x = 1; both from one assembly, so they share the pattern."""
    xyz, tets = rs.kuhn_box(n, n, n)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    t0 = np.zeros(xyz.shape[0])
    ne = tets.shape[0]
    k = rs.field_lognormal(seed, ne, 0.0, sigma) if sigma > 0 else np.ones(ne)
    q = np.zeros(ne)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    asm.assemble_heat(k, q, t0, rhoc=np.ones(ne), transient=True, alpha=1.0, dt=0.0)
    Ms = asm.matrix.to_scipy()
    asm.assemble_heat(k, q, t0)
    Ks = asm.matrix.to_scipy()
    return rs.SparseMatrix.from_scipy(Ms), rs.SparseMatrix.from_scipy(Ks), Ms, Ks


@pytest.mark.parametrize("method,ne", [(0, 4), (1, 4), (2, 1)])
def test_eigen_matches_oracle(gpu, orc, method, ne):
    rs = gpu
    _say(f"eigen method {method} ne {ne}")
    M, K, Ms, Ks = _heat_pair(rs, 5)
    econf = rs.EigenValueSolverConf(method=method, n_eigen_values=ne, n_iterations=200, solver_cvg_value=1e-12)
    mconf = rs.MatrixSolverConf(n_outer_iterations=100000, solver_cvg_value=1e-13)
    rs.seed_c_rand(7)
    d, ev = rs.REigenValueSolver(econf, mconf).solve(M, K)
    rs.seed_c_rand(7)
    d_ref, ev_ref, _ = orc.eigen_solve(Ms, Ks, method, n_eigen=ne, n_iter=200, eig_cvg=1e-12, cvg=1e-13)
    assert d.shape == d_ref.shape
    np.testing.assert_allclose(d, d_ref, rtol=1e-7)
    if method == 0:
        assert ev is None and ev_ref is None
    else:
        assert ev.shape == ev_ref.shape
        for a, b in zip(ev, ev_ref):
            assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)


def test_eigen_laplacian_first_mode(gpu):
    """k = 1, T fixed on x = 0 and x = 1, insulated elsewhere: the continuous problem's smallest
    eigenvalue is pi^2 (mode sin(pi x)); the P1 FEM value approaches it from above as h^2."""
    rs = gpu
    M, K, _, _ = _heat_pair(rs, 10, sigma=0.0)
    econf = rs.EigenValueSolverConf(method=rs.EIGEN_ARNOLDI, n_eigen_values=6, n_iterations=500,
                                    solver_cvg_value=1e-12)
    rs.seed_c_rand(1)
    d, ev = rs.REigenValueSolver(econf, rs.MatrixSolverConf(n_outer_iterations=100000, solver_cvg_value=1e-13)).solve(M, K)
    assert np.pi ** 2 <= d[0] <= 1.03 * np.pi ** 2, d
    # its vector is the sin(pi x) mode: constant along y and z up to discretisation error
    assert ev.shape == (6, K.n)


def test_eigen_invalid(gpu):
    rs = gpu
    M, K, _, _ = _heat_pair(rs, 3)
    M2, _, _, _ = _heat_pair(rs, 4)
    mconf = rs.MatrixSolverConf()
    with pytest.raises(rs.RSError):
        rs.REigenValueSolver(rs.EigenValueSolverConf(method=7), mconf).solve(M, K)
    with pytest.raises(rs.RSError):
        rs.REigenValueSolver(rs.EigenValueSolverConf(), mconf).solve(M2, K)
