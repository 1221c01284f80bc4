"""CPU pins of the post-processing restatements (oracle.heat_flux, oracle.stress_process):
closed forms for linear temperature and linear displacement fields, zero stress under rigid
motion, and node forces equal to the dense (unreduced) stiffness times the displacements."""
import numpy as np

from helpers import bary_gradients, dense_stress_K, kuhn_box_np, lame


def test_heat_flux_linear_field(orc):
    xyz, tets = kuhn_box_np(3, 2, 2, 1.0, 0.7, 0.5)
    rng = np.random.default_rng(1)
    k = np.exp(rng.standard_normal(tets.shape[0]))
    g = np.array([0.3, -1.2, 2.0])
    T = xyz @ g + 0.4
    flux = orc.heat_flux(xyz, tets, k, T)
    for e, t in enumerate(tets):
        _, V = bary_gradients(xyz[t])
        # the reference integrates grad N detJ / nInp over the points: -k detJ grad T, detJ = 6 V
        np.testing.assert_allclose(flux[e], -k[e] * 6.0 * V * g, rtol=1e-11, atol=1e-13)


def test_stress_uniform_strain_and_rigid_motion(orc):
    xyz, tets = kuhn_box_np(2, 2, 3, 1.0, 1.0, 1.5)
    ne = tets.shape[0]
    rng = np.random.default_rng(2)
    E = 1.0 + rng.random(ne)
    nu = 0.2 + 0.2 * rng.random(ne)
    G = rng.standard_normal((3, 3)) * 1e-3
    u = xyz @ G.T                                    # u_i = G_ij x_j
    nrm, sh, vm, _ = orc.stress_process(xyz, tets, E, nu, u)
    eps = 0.5 * (G + G.T)
    for e in range(ne):
        lam, mu = lame(E[e], nu[e])
        sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
        sx, sy, sz = sig[0, 0], sig[1, 1], sig[2, 2]
        n_ref = np.sqrt(sx * sx + sy * sy + sz * sz - (sx * sy + sy * sz + sz * sx))
        s_ref = np.sqrt(3 * (sig[1, 2] ** 2 + sig[0, 2] ** 2 + sig[0, 1] ** 2))
        np.testing.assert_allclose([nrm[e], sh[e], vm[e]], [n_ref, s_ref, n_ref + s_ref], rtol=1e-9)
    # rigid motion: translation + small rotation (skew G) gives no stress
    W = np.array([[0.0, -0.3, 0.2], [0.3, 0.0, -0.1], [-0.2, 0.1, 0.0]])
    u_rigid = xyz @ W.T + np.array([1.0, -2.0, 0.5])
    nrm, sh, vm, f = orc.stress_process(xyz, tets, E, nu, u_rigid)
    assert np.abs(vm).max() < 1e-12 and np.abs(f).max() < 1e-12


def test_stress_node_forces_are_K_u(orc):
    xyz, tets = kuhn_box_np(2, 2, 2, 1.0, 1.0, 1.0)
    ne = tets.shape[0]
    rng = np.random.default_rng(3)
    E = np.exp(rng.standard_normal(ne))
    nu = np.full(ne, 0.3)
    u = rng.standard_normal(xyz.shape)
    _, _, _, f = orc.stress_process(xyz, tets, E, nu, u)
    K = dense_stress_K(xyz, tets, E, nu)
    ref = (K @ u.reshape(-1)).reshape(-1, 3)
    scale = np.abs(K) @ np.abs(u.reshape(-1))
    assert np.abs(f.reshape(-1) - ref.reshape(-1)).max() <= 1e-12 * scale.max()
