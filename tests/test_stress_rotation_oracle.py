"""CPU pins of the local rotations of RSolverStress in the oracle (applyLocalRotations,
src/rsolverstress.cpp:1774-1833, and setDisplacement, :800-836):

* a skew roller -- nodes whose local first axis is a unit normal n and whose first local
  component is fixed -- built with Householder reflections (symmetric, orthogonal, so the
  reference's rotation of the load by T rather than T^T does not matter) gives the displacement
  of an independent constrained solve (null-space method on the dense C_ijkl stiffness): u = 0
  on the clamp, n . u = 0 on the roller;
* with a general rotation the load is rotated as T f (the reference's choice), not T^T f;
* the write-back returns iR^T du with iR = R^-1 on rotated nodes and zero on disabled components.
"""
import numpy as np
import scipy.sparse.linalg as spla

from helpers import bary_gradients, dense_stress_K, kuhn_box_np

EOD = 0xFFFFFFFF


def householder(n):
    """Symmetric orthogonal H with H e0 = n."""
    e0 = np.array([1.0, 0.0, 0.0])
    v = e0 - n
    if np.linalg.norm(v) < 1e-14:
        return np.eye(3)
    v /= np.linalg.norm(v)
    return np.eye(3) - 2.0 * np.outer(v, v)


def gravity_load(xyz, tets, rho, g):
    f = np.zeros(3 * xyz.shape[0])
    for e, t in enumerate(tets):
        _, V = bary_gradients(xyz[t])
        for c in range(3):
            f[3 * t + c] += rho[e] * g[c] * V / 4.0
    return f


def test_skew_roller_matches_constrained_solve(orc):
    L = 2.0
    xyz, tets = kuhn_box_np(4, 2, 2, L, 1.0, 1.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(40)
    E, nu, rho = rng.uniform(1.0, 3.0, ne), rng.uniform(0.2, 0.35, ne), np.ones(ne)
    g = np.array([0.2, -0.3, -1.0])
    clamp = np.abs(xyz[:, 0]) < 1e-12
    roller = np.abs(xyz[:, 0] - L) < 1e-12
    n = np.array([1.0, 0.6, -0.4])
    n /= np.linalg.norm(n)
    H = householder(n)
    assert np.allclose(H[:, 0], n) and np.allclose(H, H.T)
    rot = np.tile(np.eye(3), (nn, 1, 1))
    rot[roller] = H
    active = roller.astype(np.uint8)
    fixed = np.zeros((nn, 3), np.uint8)
    fixed[clamp, :] = 1
    fixed[roller, 0] = 1
    book, n_dof = orc.node_book(nn, tets, 3, fixed.reshape(-1))
    A, _, b = orc.stress_assemble_ex(xyz, tets, E, nu, rho, g, np.zeros(3 * nn), book, n_dof, rot=rot,
                                     rot_active=active)
    x = spla.spsolve(A.tocsc(), b)
    u = orc.stress_set_displacement(book, x, rot=rot, rot_active=active).reshape(nn, 3)
    # independent: minimise 1/2 u'Ku - f'u over the displacements that keep the constraints
    K = dense_stress_K(xyz, tets, E, nu)
    f = gravity_load(xyz, tets, rho, g)
    cols = []
    t1 = np.cross(n, [0.0, 0.0, 1.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    for i in range(nn):
        if clamp[i]:
            continue
        dirs = (t1, t2) if roller[i] else np.eye(3)
        for d in dirs:
            z = np.zeros(3 * nn)
            z[3 * i:3 * i + 3] = d
            cols.append(z)
    Z = np.array(cols).T
    y = np.linalg.solve(Z.T @ K @ Z, Z.T @ f)
    u_ref = (Z @ y).reshape(nn, 3)
    np.testing.assert_allclose(u, u_ref, rtol=0, atol=1e-10 * np.abs(u_ref).max())
    assert np.abs(u[roller] @ n).max() <= 1e-12 * np.abs(u_ref).max()
    assert np.abs(u[clamp]).max() == 0.0


def test_general_rotation_rotates_the_load_by_T(orc):
    xyz, tets = kuhn_box_np(2, 2, 2)
    nn, ne = xyz.shape[0], tets.shape[0]
    rng = np.random.default_rng(41)
    E, nu, rho = rng.uniform(1.0, 3.0, ne), rng.uniform(0.2, 0.35, ne), np.ones(ne)
    g = np.array([0.5, -0.2, -1.0])
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    rot = np.tile(np.eye(3), (nn, 1, 1))
    active = np.zeros(nn, np.uint8)
    sel = rng.random(nn) < 0.5
    rot[sel] = Q
    active[sel] = 1
    book, n_dof = orc.node_book(nn, tets, 3, np.zeros(3 * nn, np.uint8))
    A, _, b = orc.stress_assemble_ex(xyz, tets, E, nu, rho, g, np.zeros(3 * nn), book, n_dof, rot=rot,
                                     rot_active=active)
    T = np.zeros((3 * nn, 3 * nn))
    for i in range(nn):
        T[3 * i:3 * i + 3, 3 * i:3 * i + 3] = rot[i]
    K = dense_stress_K(xyz, tets, E, nu)
    f = gravity_load(xyz, tets, rho, g)
    np.testing.assert_allclose(A.toarray(), T.T @ K @ T, rtol=0, atol=1e-12 * np.abs(K).max())
    np.testing.assert_allclose(b, T @ f, rtol=0, atol=1e-13 * np.abs(f).max())
    assert np.abs(b - T.T @ f).max() > 1e-6 * np.abs(f).max()


def test_set_displacement_rotates_back(orc):
    rng = np.random.default_rng(42)
    nn = 12
    fixed = (rng.random(3 * nn) < 0.3).astype(np.uint8)
    book = np.full(3 * nn, EOD, np.uint32)
    book[fixed == 0] = np.arange(int((fixed == 0).sum()), dtype=np.uint32)
    x = rng.standard_normal(int((fixed == 0).sum()))
    rot = rng.standard_normal((nn, 3, 3)) + 3 * np.eye(3)
    active = (rng.random(nn) < 0.5).astype(np.uint8)
    u = orc.stress_set_displacement(book, x, rot=rot, rot_active=active).reshape(nn, 3)
    du = np.zeros(3 * nn)
    du[fixed == 0] = x
    du = du.reshape(nn, 3)
    for i in range(nn):
        expect = np.linalg.inv(rot[i]).T @ du[i] if active[i] else du[i]
        np.testing.assert_allclose(u[i], expect, rtol=1e-13, atol=1e-15)
