"""Shared test helpers: problem builders and independent numpy formulations used as pins."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

TOL_X = 1e-12


def kuhn_box_np(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    """Independent numpy Kuhn split (6 tets per cell sharing the main diagonal)."""
    px, py, pz = nx + 1, ny + 1, nz + 1
    k, j, i = np.meshgrid(np.arange(pz), np.arange(py), np.arange(px), indexing="ij")
    xyz = np.stack([i.ravel() * lx / nx, j.ravel() * ly / ny, k.ravel() * lz / nz], axis=1)
    tets = []
    for ck in range(nz):
        for cj in range(ny):
            for ci in range(nx):
                def v(bits):
                    return (ci + (bits & 1)) + px * ((cj + ((bits >> 1) & 1)) + py * (ck + ((bits >> 2) & 1)))
                for a, b in [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)]:
                    t = [v(0), v(1 << a), v((1 << a) | (1 << b)), v(7)]
                    P = xyz[t]
                    if np.linalg.det((P[1:] - P[0]).T) < 0:
                        t[2], t[3] = t[3], t[2]
                    tets.append(t)
    return xyz, np.array(tets, dtype=np.uint32)


def bary_gradients(P):
    """Gradients of the barycentric coordinates of tet P (4x3) from the inverse of the 4x4
    affine matrix [1 x y z]: a different route from the Jacobian used by the oracle."""
    M = np.hstack([np.ones((4, 1)), P])
    Minv = np.linalg.inv(M)          # columns: coefficients of each barycentric coordinate
    vol = abs(np.linalg.det(M)) / 6.0
    return Minv[1:, :].T, vol        # (4 x 3) gradients, volume


def dense_heat(xyz, tets, k, q, t_node, book, n_dof, rhoc=None, transient=False, alpha=1.0, dt=1.0):
    """Dense global assembly (all nodes), then Dirichlet elimination by partitioning."""
    nn = xyz.shape[0]
    K = np.zeros((nn, nn))
    M = np.zeros((nn, nn))
    f = np.zeros(nn)
    for e, t in enumerate(tets):
        g, V = bary_gradients(xyz[t])
        K[np.ix_(t, t)] += k[e] * V * (g @ g.T)
        f[t] += q[e] * V / 4.0
        if transient:
            M[np.ix_(t, t)] += rhoc[e] * V / 20.0 * (np.ones((4, 4)) + np.eye(4))
    if transient:
        A = M + alpha * dt * K
        rhs = dt * f + (M - (1 - alpha) * dt * K) @ t_node
    else:
        A, rhs = K, f
    free = book != 0xFFFFFFFF
    fixed = ~free
    order = np.argsort(book[free])
    fidx = np.nonzero(free)[0][order]
    Aff = A[np.ix_(fidx, fidx)]
    bf = rhs[fidx] - A[np.ix_(fidx, np.nonzero(fixed)[0])] @ t_node[fixed]
    return Aff, bf, K, f


def lame(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return lam, mu


def dense_stress_K(xyz, tets, E, nu):
    """Dense 3-dof stiffness from the isotropic elasticity tensor C_ikjl (no Voigt matrices)."""
    nn = xyz.shape[0]
    K = np.zeros((3 * nn, 3 * nn))
    I = np.eye(3)
    for e, t in enumerate(tets):
        g, V = bary_gradients(xyz[t])
        lam, mu = lame(E[e], nu[e])
        C = (lam * np.einsum("ik,jl->ikjl", I, I) + mu * (np.einsum("ij,kl->ikjl", I, I)
                                                           + np.einsum("il,kj->ikjl", I, I)))
        # K[(m,i),(n,j)] = V sum_{k,l} g_m,k C_ikjl g_n,l
        Ke = V * np.einsum("mk,ikjl,nl->minj", g, C, g).reshape(12, 12)
        dofs = (3 * np.repeat(t, 3) + np.tile(np.arange(3), 4)).astype(np.int64)
        K[np.ix_(dofs, dofs)] += Ke
    return K


def heat_problem(n, sigma=0.0, seed_k=2, seed_q=1, rs=None):
    """Unit cube, n^3 cells, T=0 at x=0 and T=1 at x=1, Q~U(0,1), k lognormal(0, sigma)."""
    xyz, tets = rs.kuhn_box(n, n, n)
    ne = tets.shape[0]
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    t_node = np.where(np.abs(x - 1) < 1e-12, 1.0, 0.0)
    k = np.ones(ne) if sigma == 0 else rs.field_lognormal(seed_k, ne, 0.0, sigma)
    q = rs.field_uniform(seed_q, ne)
    return xyz, tets, fixed, t_node, k, q


def random_spd(n, density, seed, diag_shift=1.0):
    rng = np.random.default_rng(seed)
    R = sp.random(n, n, density=density, random_state=rng, format="csr")
    A = R @ R.T + sp.eye(n) * diag_shift
    return A.tocsr()


def csr_arrays(A):
    A = A.tocsr()
    A.sort_indices()
    return A.shape[0], A.indptr.astype(np.uint32), A.indices.astype(np.uint32), A.data.astype(np.float64)
