"""Synthetic radiative-heat workload: the six inner faces of the unit box, each cut into k x k
square patches of two triangles, with view factors from the differential-area formula
F_ij = cos(t_i) cos(t_j) A_j / (pi r^2) between patch centres (zero within a face), scaled
symmetrically so that the enclosure is closed (rows sum to 1) and reciprocal (A_i F_ij = A_j F_ji).
This stands in for a view-factor matrix read from file or computed by RHemiCube, which is outside
the solve path."""
from __future__ import annotations

import numpy as np


def box_enclosure(k: int):
    """Returns (xyz, tris, patch_ptr, patch_elems, F, patch_area, patch_face).  Patch p owns
    elements 2p and 2p + 1; faces in the order x=0, x=1, y=0, y=1, z=0, z=1."""
    g = np.linspace(0.0, 1.0, k + 1)
    xyz, tris, centres, normals = [], [], [], []
    node0 = 0
    for axis in range(3):
        for side in (0.0, 1.0):
            u_ax, v_ax = [a for a in range(3) if a != axis]
            uu, vv = np.meshgrid(g, g, indexing="ij")
            pts = np.zeros(((k + 1) * (k + 1), 3))
            pts[:, axis] = side
            pts[:, u_ax] = uu.ravel()
            pts[:, v_ax] = vv.ravel()
            xyz.append(pts)
            n = np.zeros(3)
            n[axis] = 1.0 if side == 0.0 else -1.0          # inward normal
            for i in range(k):
                for j in range(k):
                    a = node0 + i * (k + 1) + j
                    b, c, d = a + (k + 1), a + (k + 1) + 1, a + 1
                    tris.append((a, b, c))
                    tris.append((a, c, d))
                    ctr = np.zeros(3)
                    ctr[axis] = side
                    ctr[u_ax] = (g[i] + g[i + 1]) / 2
                    ctr[v_ax] = (g[j] + g[j + 1]) / 2
                    centres.append(ctr)
                    normals.append(n)
            node0 += (k + 1) * (k + 1)
    xyz = np.concatenate(xyz)
    tris = np.array(tris, np.uint32)
    centres = np.array(centres)
    normals = np.array(normals)
    npch = centres.shape[0]
    area = np.full(npch, 1.0 / (k * k))
    face = np.repeat(np.arange(6), k * k)
    F = np.zeros((npch, npch))
    step = 1024
    for r0 in range(0, npch, step):
        r1 = min(npch, r0 + step)
        d = centres[None, :, :] - centres[r0:r1, None, :]           # i -> j
        r2 = np.einsum("ijk,ijk->ij", d, d)
        r2[r2 == 0] = 1.0
        ci = np.einsum("ijk,ik->ij", d, normals[r0:r1])
        cj = -np.einsum("ijk,jk->ij", d, normals)
        Fb = np.clip(ci, 0, None) * np.clip(cj, 0, None) / (np.pi * r2 * r2) * area[None, :]
        Fb[face[r0:r1, None] == face[None, :]] = 0.0
        F[r0:r1] = Fb
    # the point formula over-counts near edges and corners: scale G = A_i F_ij symmetrically
    # (G <- D G D) until every row of F sums to 1, which keeps the reciprocity
    G = F * area[:, None]
    for _ in range(30):
        d = np.sqrt(area / G.sum(1))
        G *= d[:, None]
        G *= d[None, :]
    F = G / area[:, None]
    patch_ptr = np.arange(0, 2 * npch + 1, 2, dtype=np.uint32)
    patch_elems = np.arange(2 * npch, dtype=np.uint32)
    return xyz, tris, patch_ptr, patch_elems, F, area, face
