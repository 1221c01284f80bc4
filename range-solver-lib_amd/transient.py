"""Batched transient heat problems on the GPU assembly, solved by the batched Jacobi-CG.

System i: the unit cube in k_i^3 cells, each split into 6 tets (Kuhn split); per-tet
conductivity lognormal(0, sigma) and source U(0, 1); rho*c = 1; T = 0 on x = 0 and T = 1 on
x = 1, T = 0 elsewhere at t = 0; implicit Euler (alpha = 1) with step dt.

RSolverHeat assembles every step with the previous temperatures (src/rsolverheat.cpp:757-800):
A = M + dt K and b_k = dt f + M T_{k-1} minus the Dirichlet lifting.  With fixed coefficients
A does not change, and b_k = c + M_ff x_{k-1} where c is b assembled with the free
temperatures at zero.  The builder therefore assembles twice per system on the GPU: once with
dt = 0 (A = M, the mass matrix) and once with the step dt (A and c), and hands both to the
batch, which forms every later right-hand side on the device.
"""
from __future__ import annotations

import time

import numpy as np


def system_sizes(rs, count: int, kmin: int, kmax: int, seed: int) -> np.ndarray:
    """Cells per edge of each system, uniform on {kmin..kmax} from the library's seeded stream."""
    u = rs.field_uniform(seed, count)
    span = kmax - kmin + 1
    return (kmin + np.minimum((u * span).astype(np.int64), span - 1)).astype(np.int64)


def field_seeds(seed: int, i: int) -> tuple[int, int]:
    """Seeds of system i's conductivity and source fields."""
    return 1_000_003 * seed + 2 * i, 1_000_003 * seed + 2 * i + 1


class Mesh:
    """One cube mesh and its GPU assembly (shared by every system of that size)."""

    def __init__(self, rs, k: int):
        self.k = k
        self.xyz, self.tets = rs.kuhn_box(k, k, k)
        x = self.xyz[:, 0]
        self.fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
        self.t0 = np.where(np.abs(x - 1) < 1e-12, 1.0, 0.0)
        self.asm = rs.Assembly(self.xyz, self.tets, 1, self.fixed)
        self.n = self.asm.n_dof
        self.nnz = self.asm.nnz
        self.dT = rs.DeviceArray.from_host(self.t0)
        self.drhoc = rs.DeviceArray.from_host(np.ones(self.tets.shape[0]))


def fields(rs, mesh: Mesh, seed: int, i: int, sigma: float):
    sk, sq = field_seeds(seed, i)
    ne = mesh.tets.shape[0]
    return rs.field_lognormal(sk, ne, 0.0, sigma), rs.field_uniform(sq, ne)


def assemble_system(rs, B, i: int, m: Mesh, dk, dq, dc, dt: float) -> None:
    """System i of batch B from its resident fields: the mass matrix (the transient assembly
    with dt = 0), then A = M + dt K and c (the assembly with the step dt)."""
    m.asm.assemble_heat_device(dk, dq, m.dT, dc, drhoc=m.drhoc, transient=True, alpha=1.0, dt=0.0)
    B.set_mass(i, m.asm.matrix)
    m.asm.assemble_heat_device(dk, dq, m.dT, dc, drhoc=m.drhoc, transient=True, alpha=1.0, dt=dt)
    B.set_matrix(i, m.asm.matrix)
    B.set_rhs(i, dc)


def build_batch(rs, sizes, sigma: float = 1.5, dt: float = 1e-3, seed: int = 4, log=None):
    """Assemble every system on the GPU and load it into one rs.Batch.  Returns
    (batch, meshes by size, build seconds, seconds of the GPU assembly alone).  The meshes,
    their patterns and the seeded fields are prepared and uploaded first; the second figure
    times only the assembly of every system into the batch from those resident inputs."""
    t0 = time.time()
    sizes = [int(k) for k in sizes]
    meshes = {}
    for k in sorted(set(sizes)):
        meshes[k] = Mesh(rs, k)
    n_rows = [meshes[k].n for k in sizes]
    nnz = [meshes[k].nnz for k in sizes]
    B = rs.Batch(n_rows, nnz)
    dev = []
    for i, k in enumerate(sizes):
        kf, qf = fields(rs, meshes[k], seed, i, sigma)
        dev.append((rs.DeviceArray.from_host(kf), rs.DeviceArray.from_host(qf), rs.DeviceArray(meshes[k].n)))
    rs.synchronize()
    ta = time.time()
    for i, k in enumerate(sizes):
        dk, dq, dc = dev[i]
        assemble_system(rs, B, i, meshes[k], dk, dq, dc, dt)
        if log and (i + 1) % 128 == 0:
            log(f"assembled {i + 1}/{len(sizes)} systems, {time.time() - ta:.2f}s")
    rs.synchronize()
    t_asm = time.time() - ta
    return B, meshes, time.time() - t0, t_asm
