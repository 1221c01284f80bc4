"""GPU parity tests of the coloured tet assembly (heat, stress) and the assembled solves.

The CSR pattern is compared bit-exactly with the oracle's RSparseMatrix restatement; values
and right-hand sides to 1e-12 relative (element matrices are summed in colour order on the
GPU and in element order by the oracle).
"""
import numpy as np
import pytest

from helpers import heat_problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _cmp_csr(A_gpu, A_ref, rtol=1e-12):
    A_gpu.sort_indices()
    A_ref.sort_indices()
    assert A_gpu.shape == A_ref.shape
    assert np.array_equal(A_gpu.indptr, A_ref.indptr)
    assert np.array_equal(A_gpu.indices, A_ref.indices)
    scale = abs(A_ref.data).max()
    assert abs(A_gpu.data - A_ref.data).max() <= rtol * scale


@pytest.mark.parametrize("n,sigma", [(4, 0.0), (20, 1.0)])
def test_heat_assembly_matches_oracle(gpu, orc, n, sigma):
    _say(f"heat assembly n={n}")
    rs = gpu
    xyz, tets, fixed, t_node, k, q = heat_problem(n, sigma, rs=rs)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    b = asm.assemble_heat(k, q, t_node)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    assert np.array_equal(asm.book(), book) and asm.n_dof == n_dof
    A_ref, b_ref = orc.heat_assemble(xyz, tets, k, q, t_node, book, n_dof)
    _cmp_csr(asm.matrix.to_scipy(), A_ref)
    np.testing.assert_allclose(b, b_ref, rtol=1e-12, atol=1e-13 * abs(b_ref).max())
    # deterministic: a second assembly is bit-identical
    b2 = asm.assemble_heat(k, q, t_node)
    assert np.array_equal(b, b2)
    assert np.array_equal(asm.matrix.download()[2], asm.matrix.download()[2])
    assert asm.n_colours <= 64


@pytest.mark.parametrize("transient", [False, True])
def test_heat_row_gather_equals_colour_launches(gpu, transient):
    """Both heat kernels add every entry's contributions in colour-sorted element order."""
    rs = gpu
    xyz, tets, fixed, t_node, k, q = heat_problem(14, 2.0, rs=rs)
    rng = np.random.default_rng(5)
    rhoc = 1.0 + rng.random(tets.shape[0])
    asm = rs.Assembly(xyz, tets, 1, fixed)
    kw = dict(rhoc=rhoc, transient=True, alpha=0.7, dt=1e-3) if transient else {}
    b_rows = asm.assemble_heat(k, q, t_node, **kw)
    v_rows = asm.matrix.download()[2]
    asm.set_method(False)
    b_col = asm.assemble_heat(k, q, t_node, **kw)
    v_col = asm.matrix.download()[2]
    scale = abs(v_col).max()
    assert abs(v_rows - v_col).max() <= 1e-14 * scale
    assert abs(b_rows - b_col).max() <= 1e-14 * abs(b_col).max()


def test_heat_transient_assembly_matches_oracle(gpu, orc):
    rs = gpu
    xyz, tets, fixed, t_node, k, q = heat_problem(6, 1.5, rs=rs)
    rng = np.random.default_rng(3)
    rhoc = 1.0 + rng.random(tets.shape[0])
    t_old = t_node + 0.2 * rng.random(t_node.size) * (fixed == 0)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    b = asm.assemble_heat(k, q, t_old, rhoc=rhoc, transient=True, alpha=1.0, dt=1e-3)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    A_ref, b_ref = orc.heat_assemble(xyz, tets, k, q, t_old, book, n_dof, rhoc=rhoc, transient=True,
                                     alpha=1.0, dt=1e-3)
    _cmp_csr(asm.matrix.to_scipy(), A_ref)
    np.testing.assert_allclose(b, b_ref, rtol=1e-11, atol=1e-13 * abs(b_ref).max())


def test_heat_config_20_cube_solve(gpu, orc):
    """The 20^3 heat configuration: k=1, Q~U(0,1) (seed 1), T=0/1 at x=0/1, FP64 Jacobi-CG to
    the reference's criterion 1e-10; GPU assembly + solve vs oracle assembly + solve."""
    rs = gpu
    xyz, tets, fixed, t_node, k, q = heat_problem(20, 0.0, rs=rs)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    assert (asm.n_dof, asm.nnz) == (8379, 116015)
    b = asm.assemble_heat(k, q, t_node)
    S = rs.RMatrixSolver(rs.MatrixSolverConf(n_outer_iterations=10000, solver_cvg_value=1e-10))
    x = S.solve(asm.matrix, b, None, rs.PRECOND_JACOBI, 1)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    A_ref, b_ref = orc.heat_assemble(xyz, tets, k, q, t_node, book, n_dof)
    x_ref, info_ref = orc.solve(A_ref, b_ref, precond=orc.PRECOND_JACOBI, n_outer=10000, cvg=1e-10)
    _say(f"C1: gpu its={S.info.iterations} oracle its={info_ref['iterations']}")
    assert abs(S.info.iterations - info_ref["iterations"]) <= 1
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-8
    # maximum principle-ish sanity: with Q >= 0 the temperature exceeds the linear profile
    free = book != rs.RS_EOD
    assert np.all(x[book[free]] >= xyz[free, 0] - 1e-9)


def _cantilever(rs, nx, ny, nz, L=4.0, sigma=1.0, seed=3):
    xyz, tets = rs.kuhn_box(nx, ny, nz, L, 1.0, 1.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    clamp = np.abs(xyz[:, 0]) < 1e-12
    fixed = np.repeat(clamp, 3).astype(np.uint8)
    E = rs.field_lognormal(seed, ne, 0.0, sigma)
    nu = np.full(ne, 0.3)
    rho = np.ones(ne)
    tris = rs.plane_faces(xyz, tets, 0, L)
    trac = np.tile([0.0, 0.0, -1.0], (tris.shape[0], 1))
    return xyz, tets, fixed, E, nu, rho, tris, trac


def test_stress_assembly_matches_oracle(gpu, orc):
    rs = gpu
    xyz, tets, fixed, E, nu, rho, tris, trac = _cantilever(rs, 8, 3, 3)
    nn = xyz.shape[0]
    u_node = np.zeros(3 * nn)
    grav = np.array([0.0, 0.0, -0.5])
    asm = rs.Assembly(xyz, tets, 3, fixed)
    b = asm.assemble_stress(E, nu, rho, grav, u_node, tris, trac)
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    assert np.array_equal(asm.book(), book)
    A_ref, b_ref = orc.stress_assemble(xyz, tets, E, nu, rho, grav, u_node, tris, trac, book, n_dof)
    _cmp_csr(asm.matrix.to_scipy(), A_ref, rtol=1e-12)
    np.testing.assert_allclose(b, b_ref, rtol=1e-11, atol=1e-13 * abs(b_ref).max())
    # block-Jacobi(3) CG on the assembled system vs the oracle
    S = rs.RMatrixSolver(rs.MatrixSolverConf(n_outer_iterations=20000, solver_cvg_value=1e-12))
    x = S.solve(asm.matrix, b, None, rs.PRECOND_BLOCK_JACOBI, 3)
    x_ref, info_ref = orc.solve(A_ref, b_ref, precond=2, block_size=3, n_outer=20000, cvg=1e-12)
    assert abs(S.info.iterations - info_ref["iterations"]) <= 1
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-8


def test_stress_prescribed_displacement_lifting(gpu, orc):
    rs = gpu
    xyz, tets, fixed, E, nu, rho, tris, trac = _cantilever(rs, 4, 2, 2)
    nn = xyz.shape[0]
    u_node = np.zeros(3 * nn)
    clamp = np.abs(xyz[:, 0]) < 1e-12
    u_node[0::3][clamp] = 1e-3 * xyz[clamp, 1]   # nonzero prescribed values
    asm = rs.Assembly(xyz, tets, 3, fixed)
    b = asm.assemble_stress(E, nu, rho, [0, 0, 0], u_node)
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    _, b_ref = orc.stress_assemble(xyz, tets, E, nu, rho, [0, 0, 0], u_node, None, None, book, n_dof)
    np.testing.assert_allclose(b, b_ref, rtol=1e-11, atol=1e-13 * abs(b_ref).max())


def test_dofs_to_nodes(gpu):
    rs = gpu
    xyz, tets, fixed, t_node, k, q = heat_problem(5, 0.0, rs=rs)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    x = np.arange(asm.n_dof, dtype=np.float64)
    dx = rs.DeviceArray.from_host(x)
    dn = rs.DeviceArray.from_host(t_node)
    asm.dofs_to_nodes(dx, dn)
    out = dn.download()
    book = asm.book()
    free = book != rs.RS_EOD
    assert np.array_equal(out[free], x[book[free]])
    assert np.array_equal(out[~free], t_node[~free])
