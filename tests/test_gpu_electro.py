"""GPU parity tests of the electrostatics path (rs_assemble_electrostatics_device,
rs_electrostatics_process_device) against the oracle, and the parallel-plate patch test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plates(rs, n, volts=5.0):
    xyz, tets = rs.kuhn_box(n, n, n)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    phi = np.where(np.abs(x - 1) < 1e-12, volts, 0.0)
    return xyz, tets, fixed, phi


def _cmp_csr(A_gpu, A_ref, rtol=1e-12):
    A_gpu.sort_indices()
    A_ref.sort_indices()
    assert np.array_equal(A_gpu.indptr, A_ref.indptr) and np.array_equal(A_gpu.indices, A_ref.indices)
    assert abs(A_gpu.data - A_ref.data).max() <= rtol * abs(A_ref.data).max()


def test_electrostatics_matches_oracle(gpu, orc):
    rs = gpu
    xyz, tets, fixed, phi = _plates(rs, 8)
    ne = tets.shape[0]
    eps_r = rs.field_lognormal(7, ne, 0.0, 1.0)
    rho = (rs.field_uniform(8, ne) - 0.5) * 1e-9
    sigma = rs.field_lognormal(9, ne, 2.0, 1.0)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    db = rs.DeviceArray(asm.n_dof)
    deps, drho, dphi = (rs.DeviceArray.from_host(a) for a in (eps_r, rho, phi))
    asm.assemble_electrostatics_device(deps, drho, dphi, db)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    A_ref, b_ref = orc.electrostatics_assemble(xyz, tets, eps_r, rho, phi, book, n_dof)
    _cmp_csr(asm.matrix.to_scipy(), A_ref)
    b = db.download()
    assert np.abs(b - b_ref).max() <= 1e-12 * np.abs(b_ref).max()
    # processing of an arbitrary potential field
    phi_any = np.random.default_rng(3).random(xyz.shape[0]) * 5.0
    outs = [rs.DeviceArray(3 * ne), rs.DeviceArray(3 * ne), rs.DeviceArray(ne), rs.DeviceArray(ne), rs.DeviceArray(ne)]
    asm.electrostatics_process_device(deps, rs.DeviceArray.from_host(sigma), rs.DeviceArray.from_host(phi_any), *outs)
    ref = orc.electrostatics_process(xyz, tets, eps_r, sigma, phi_any)
    names = ("field", "current", "energy", "resistivity", "joule")
    for o, r, name in zip(outs, ref, names):
        g = o.download().reshape(r.shape)
        assert np.abs(g - r).max() <= 1e-11 * np.abs(r).max(), name


def test_parallel_plates_patch(gpu):
    """Uniform permittivity, no charge, 0 V / 5 V plates: the potential is exactly linear and the
    field uniform (the reference's B carries detJ, so E = -detJ * 5 x-hat per tet)."""
    rs = gpu
    xyz, tets, fixed, phi = _plates(rs, 6)
    ne = tets.shape[0]
    asm = rs.Assembly(xyz, tets, 1, fixed)
    deps, drho, dphi = rs.DeviceArray.from_host(np.full(ne, 2.5)), rs.DeviceArray.from_host(np.zeros(ne)), \
        rs.DeviceArray.from_host(phi)
    db, dx = rs.DeviceArray(asm.n_dof), rs.DeviceArray.from_host(np.zeros(asm.n_dof))
    asm.assemble_electrostatics_device(deps, drho, dphi, db)
    conf = rs.MatrixSolverConf(n_outer_iterations=10000, solver_cvg_value=1e-12, stop_rule=rs.STOP_RELRES)
    info = rs.RMatrixSolver(conf).solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
    assert info.converged == 1
    asm.dofs_to_nodes(dx, dphi)
    phi_n = dphi.download()
    np.testing.assert_allclose(phi_n, 5.0 * xyz[:, 0], atol=1e-9)
    outs = [rs.DeviceArray(3 * ne), rs.DeviceArray(3 * ne), rs.DeviceArray(ne), rs.DeviceArray(ne), rs.DeviceArray(ne)]
    asm.electrostatics_process_device(deps, rs.DeviceArray.from_host(np.ones(ne)), dphi, *outs)
    field = outs[0].download().reshape(ne, 3)
    V = 1.0 / (6 * 6 ** 3)
    np.testing.assert_allclose(field, np.tile([-5.0 * 6 * V, 0.0, 0.0], (ne, 1)), atol=1e-9)


def test_electrostatics_multilevel_against_oracle(gpu, orc):
    """The electrostatics system through the multilevel cycle (node bins, exact coarsest inverse)
    in mixed-precision flexible CG: the oracle's direct solution of its own assembly to 1e-6, in
    a fraction of Jacobi-CG's iterations on the same system (no counterpart in the reference)."""
    import scipy.sparse.linalg as spla
    rs = gpu
    c = 24
    xyz, tets, fixed, phi = _plates(rs, c)
    ne = tets.shape[0]
    eps_r = rs.field_lognormal(7, ne, 0.0, 1.0)
    rho = (rs.field_uniform(8, ne) - 0.5) * 1e-9
    asm = rs.Assembly(xyz, tets, 1, fixed)
    n = asm.n_dof
    db = rs.DeviceArray(n)
    asm.assemble_electrostatics_device(*(rs.DeviceArray.from_host(a) for a in (eps_r, rho, phi)), db)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    A_ref, b_ref = orc.electrostatics_assemble(xyz, tets, eps_r, rho, phi, book, n_dof)
    x_ref = spla.spsolve(A_ref.tocsc(), b_ref)
    L = rs.LearnedPreconditioner.from_bins(asm.matrix, asm, (c // 2, c // 2, c // 2), max_coarse=rs.MAX_COARSE)
    L.set_cycle(1, 1, rs.COARSE_EXACT)
    assert L.n_levels >= 2
    conf = rs.MatrixSolverConf(n_outer_iterations=5000, solver_cvg_value=1e-8, stop_rule=rs.STOP_RELRES,
                               precision=rs.PRECISION_MIXED)
    dx = rs.DeviceArray.from_host(np.zeros(n))
    info = rs.RMatrixSolver(conf).solve_learned_device(asm.matrix, L, db, dx)
    assert info.converged == 1 and info.true_relres <= 1e-8
    x = dx.download()
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= 1e-6
    dj = rs.DeviceArray.from_host(np.zeros(n))
    jac = rs.RMatrixSolver(rs.MatrixSolverConf(n_outer_iterations=100000, solver_cvg_value=1e-8,
                                               stop_rule=rs.STOP_RELRES)).solve_device(asm.matrix, db, dj,
                                                                                       rs.PRECOND_JACOBI, 1)
    print(f"electrostatics {c}^3: multilevel {info.iterations} its, Jacobi-CG {jac.iterations} its", flush=True)
    assert info.iterations * 3 < jac.iterations
