"""CPU pin of the magnetostatics restatement: the assembled system against a dense formulation
(per-component negative Laplacian from barycentric gradients, source e0 V/4 (grad N_m x J_m),
elimination of the fixed nodes with the prescribed field)."""
import numpy as np

from helpers import bary_gradients, kuhn_box_np

E0 = 8.854187817e-12


def test_magnetostatics_assembly_dense(orc):
    xyz, tets = kuhn_box_np(2, 3, 2)
    nn = xyz.shape[0]
    x = xyz[:, 0]
    fixed_node = (np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)
    fixed = np.repeat(fixed_node, 3).astype(np.uint8)
    rng = np.random.default_rng(4)
    J = rng.standard_normal((nn, 3)) * 1e6
    H = np.where(fixed_node[:, None], rng.standard_normal((nn, 3)), 0.0)
    book, n_dof = orc.node_book(nn, tets, 3, fixed)
    A, b = orc.magnetostatics_assemble(xyz, tets, J, H, book, n_dof)
    K = np.zeros((3 * nn, 3 * nn))
    f = np.zeros(3 * nn)
    for t in tets:
        g, V = bary_gradients(xyz[t])
        for m in range(4):
            for n in range(4):
                for c in range(3):
                    K[3 * t[m] + c, 3 * t[n] + c] -= V * (g[m] @ g[n])
            f[3 * t[m]:3 * t[m] + 3] += E0 * V / 4 * np.cross(g[m], J[t[m]])
    free = book != 0xFFFFFFFF
    fi = np.nonzero(free)[0][np.argsort(book[free])]
    fx = np.nonzero(~free)[0]
    Ad = K[np.ix_(fi, fi)]
    bd = f[fi] - K[np.ix_(fi, fx)] @ H.reshape(-1)[fx]
    Ao = A.toarray()
    assert np.abs(Ao - Ad).max() <= 1e-12 * np.abs(Ad).max()
    np.testing.assert_allclose(b, bd, rtol=1e-9, atol=1e-12 * np.abs(bd).max())
