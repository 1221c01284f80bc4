"""CPU pins of the electrostatics restatement: the assembled system against the dense
formulation with permittivity e0 eps_r and charge -rho, and the processed quantities for a
linear potential in closed form."""
import numpy as np

from helpers import bary_gradients, dense_heat, kuhn_box_np

E0 = 8.854187817e-12


def test_electrostatics_assembly_dense(orc):
    xyz, tets = kuhn_box_np(3, 3, 2)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    phi = np.where(np.abs(x - 1) < 1e-12, 5.0, 0.0)
    rng = np.random.default_rng(1)
    eps_r = 1.0 + 4.0 * rng.random(tets.shape[0])
    rho = rng.standard_normal(tets.shape[0]) * 1e-9
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    A, b = orc.electrostatics_assemble(xyz, tets, eps_r, rho, phi, book, n_dof)
    Ad, bd, _, _ = dense_heat(xyz, tets, eps_r * E0, -rho, phi, book, n_dof)
    assert np.abs(A.toarray() - Ad).max() <= 1e-12 * np.abs(Ad).max()
    np.testing.assert_allclose(b, bd, rtol=1e-10, atol=1e-12 * np.abs(bd).max())


def test_electrostatics_process_linear_potential(orc):
    xyz, tets = kuhn_box_np(2, 2, 2, 1.0, 0.5, 0.8)
    ne = tets.shape[0]
    rng = np.random.default_rng(2)
    eps_r = 1.0 + rng.random(ne)
    sigma = 10.0 ** rng.uniform(-2, 3, ne)
    a = np.array([2.0, -1.0, 0.5])
    phi = xyz @ a + 3.0
    f, j, en, res, jh = orc.electrostatics_process(xyz, tets, eps_r, sigma, phi)
    for e, t in enumerate(tets):
        g, V = bary_gradients(xyz[t])
        B = g * 6.0 * V                               # grad N detJ, the reference's B
        E = -6.0 * V * a
        np.testing.assert_allclose(f[e], E, rtol=1e-11)
        np.testing.assert_allclose(j[e], sigma[e] * E, rtol=1e-11)
        np.testing.assert_allclose(en[e], eps_r[e] * E0 * (E @ E) / 2, rtol=1e-11)
        np.testing.assert_allclose(res[e], 1.0 / sigma[e], rtol=1e-11)
        s = E / np.linalg.norm(E)
        L = 2.0 / np.abs(B @ s).sum()
        np.testing.assert_allclose(jh[e], sigma[e] * (E @ E) * L, rtol=1e-10)
