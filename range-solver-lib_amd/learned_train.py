#!/usr/bin/env python3
"""Training of the learned coarsest-level solver on the GPU.

Data come from the solver's own runs: for each seeded heat problem (unit cube, cells^3 x 6
tets, conductivity lognormal(0, sigma), T=0/1 on x=0/1) the flexible CG with the multilevel
preconditioner records the right-hand side that reaches the coarsest level in each iteration
(rs_learned_record).  The exact coarse error of such a residual r is e* = A_c^-1 r, and
s = r.A_c^-1.r = ||e*||^2_{A_c}.  The network is trained by the library's HIP forward, backward
and Adam kernels (rs_trainer_*) to minimise the A_c-norm error fraction it leaves,
(e.A_c.e - 2 e.r)/s + 1, averaged over a batch.

Usage (writes the weights next to this file):
    python range-solver-lib_amd/learned_train.py --problems 2000 --cells 64 --sigma 2.5
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
WEIGHTS = os.path.join(HERE, "weights", "heat_coarse512.npz")


def _package():
    sys.path.insert(0, ROOT)
    import __graft_entry__
    return __graft_entry__.load_package()


def say(msg: str) -> None:
    print(f"[train] {msg}", flush=True)


def heat_fields(rs, ne: int, seed: int, sigma: float):
    return rs.field_lognormal(seed, ne, 0.0, sigma), rs.field_uniform(seed + 7919, ne)


class HeatFamily:
    """One mesh (cells^3, Kuhn split, T=0/1 on x=0/1) and its hierarchy; problems differ in
    their conductivity field."""

    def __init__(self, rs, cells: int, max_coarse: int = 512):
        self.rs = rs
        xyz, tets = rs.kuhn_box(cells, cells, cells)
        x = xyz[:, 0]
        fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
        self.xyz, self.tets = xyz, tets
        self.t_node = np.where(np.abs(x - 1) < 1e-12, 1.0, 0.0)
        self.asm = rs.Assembly(xyz, tets, 1, fixed)
        self.n = self.asm.n_dof
        self.maps = rs.hierarchy_by_bins(xyz, self.asm.book(), 1, (cells // 2,) * 3, max_coarse=max_coarse)
        self.dT = rs.DeviceArray.from_host(self.t_node)
        self.db = rs.DeviceArray(self.n)
        self.dx = rs.DeviceArray(self.n)
        self.zero = rs.DeviceArray.from_host(np.zeros(self.n))

    def assemble(self, seed: int, sigma: float):
        k, q = heat_fields(self.rs, self.tets.shape[0], seed, sigma)
        self.dk, self.dq = self.rs.DeviceArray.from_host(k), self.rs.DeviceArray.from_host(q)
        self.asm.assemble_heat_device(self.dk, self.dq, self.dT, self.db)

    def solve(self, L, tol: float = 1e-8, n_outer: int = 2000):
        rs = self.rs
        rs._check(rs.load().rs_memcpy_dtod(self.dx.ptr, self.zero.ptr, 8 * self.n))
        conf = rs.MatrixSolverConf(n_outer_iterations=n_outer, solver_cvg_value=tol, stop_rule=rs.STOP_RELRES,
                                   precision=rs.PRECISION_MIXED)
        return rs.RMatrixSolver(conf).solve_learned_device(self.asm.matrix, L, self.db, self.dx)


def coarse_pairs(L, count: int):
    """Recorded coarsest right-hand sides, the coarsest operator (CSR) and s = r.A_c^-1.r."""
    import scipy.linalg as sla
    R = L.records(count).astype(np.float64)
    C = L.level_matrix(L.n_levels - 1)
    Ad = C.toarray()
    cho = sla.cho_factor(Ad)
    E = sla.cho_solve(cho, R.T).T
    S = np.einsum("bi,bi->b", E, R)
    return C, R, S


def generate(rs, fam: HeatFamily, seeds, sigma: float, record: int, log_every: int = 50, policy=None):
    """Pairs from the solver's runs with the coarsest level solved by coarse Jacobi, or by the
    network with parameters `policy` (the trajectories a trained network actually sees)."""
    L = rs.LearnedPreconditioner(fam.asm.matrix, fam.maps, params=rs.init_params(0) if policy is None else policy)
    L.set_cycle(1, 1, rs.COARSE_JACOBI if policy is None else rs.COARSE_NETWORK)
    L.record(record)
    data, its = [], []
    t0 = time.time()
    for i, seed in enumerate(seeds):
        fam.assemble(seed, sigma)
        info = fam.solve(L)
        cnt = min(info.iterations, record)
        C, R, S = coarse_pairs(L, cnt)
        keep = S > 0
        data.append((C.data.astype(np.float32), R[keep].astype(np.float32), S[keep].astype(np.float32)))
        its.append(info.iterations)
        if (i + 1) % log_every == 0 or i + 1 == len(seeds):
            say(f"data: {i + 1}/{len(seeds)} problems, {sum(len(d[1]) for d in data)} pairs, "
                f"mean its {np.mean(its):.1f}, {time.time() - t0:.0f}s")
    return L, C.indptr.astype(np.uint32), C.indices.astype(np.uint32), data


def batches(data, batch: int, rng):
    index = np.array([(p, j) for p, d in enumerate(data) for j in range(len(d[1]))], dtype=np.int64)
    while True:
        rng.shuffle(index)
        for k in range(0, len(index) - batch + 1, batch):
            sel = index[k:k + batch]
            cv = np.stack([data[p][0] for p, _ in sel])
            R = np.stack([data[p][1][j] for p, j in sel])
            S = np.array([data[p][2][j] for p, j in sel], np.float32)
            yield cv, R, S


def evaluate(T, cptr, ccol, data, batch: int, max_batches: int = 16) -> float:
    """Mean loss over (up to max_batches batches of) the held-out pairs."""
    losses = []
    for k, (cv, R, S) in enumerate(batches_once(data, batch)):
        if k == max_batches:
            break
        losses.append(T.gradient_csr(cptr, ccol, cv, R, S)[0])
    return float(np.mean(np.concatenate(losses))) if losses else float("nan")


def batches_once(data, batch: int):
    index = [(p, j) for p, d in enumerate(data) for j in range(len(d[1]))]
    for k in range(0, len(index), batch):
        sel = index[k:k + batch]
        yield (np.stack([data[p][0] for p, _ in sel]), np.stack([data[p][1][j] for p, j in sel]),
               np.array([data[p][2][j] for p, j in sel], np.float32))


def train(rs, cells=64, sigma=2.5, problems=2000, seed0=10000, holdout=0.05, record=64, batch=64,
          steps=4000, lr=1e-3, out=WEIGHTS, init_seed=0, log_every=200, init=None, rounds=1):
    """rounds > 1: after each round the pairs are regenerated with the current network solving
    the coarsest level (on-policy data) and training continues from the current parameters."""
    say(f"family {cells}^3 sigma={sigma}: mesh, assembly set-up, hierarchy")
    fam = HeatFamily(rs, cells)
    say(f"unknowns {fam.n}, levels {[m.size for m in fam.maps]} -> {int(fam.maps[-1].max()) + 1}")
    seeds = list(range(seed0, seed0 + problems))
    n_hold = max(1, int(round(holdout * problems)))
    theta = rs.init_params(init_seed) if init is None else np.asarray(init, np.float32)
    for rnd in range(rounds):
        policy = None if (init is None and rnd == 0) else theta
        say(f"round {rnd + 1}/{rounds}: pairs from {'coarse Jacobi' if policy is None else 'the network'} runs")
        theta, summary = _train_round(rs, fam, seeds, n_hold, sigma, record, batch, steps, lr, out, init_seed,
                                      log_every, theta, policy)
        summary.update({"cells": cells, "sigma": sigma, "problems": problems, "seed0": seed0, "round": rnd + 1,
                        "rounds": rounds})
        if out:
            np.savez(out, params=theta.astype(np.float32), summary=json.dumps(summary))
            say(f"wrote {out}")
    return theta, summary


def _train_round(rs, fam, seeds, n_hold, sigma, record, batch, steps, lr, out, init_seed, log_every, theta0, policy):
    L, cptr, ccol, data = generate(rs, fam, seeds, sigma, record, policy=policy)
    train_data, hold_data = data[:-n_hold], data[-n_hold:]
    T = rs.Trainer(theta0, max_batch=batch, max_nodes=int(cptr.size - 1), max_nnz=int(ccol.size))
    base = evaluate(T, cptr, ccol, hold_data, batch)
    say(f"held-out loss before this round: {base:.4f}")
    rng = np.random.default_rng(init_seed)
    gen = batches(train_data, batch, rng)
    curve = []
    t0 = time.time()
    for step in range(1, steps + 1):
        a = lr * 0.5 * (1.0 + np.cos(np.pi * (step - 1) / steps))
        cv, R, S = next(gen)
        m = T.step_csr(cptr, ccol, cv, R, S, a)
        if step % log_every == 0 or step == steps:
            held = evaluate(T, cptr, ccol, hold_data, batch)
            curve.append((step, float(m), held))
            say(f"step {step}/{steps}: batch loss {m:.4f}, held-out {held:.4f}, lr {a:.2e}, {time.time() - t0:.0f}s")
            if out:
                os.makedirs(os.path.dirname(out), exist_ok=True)
                np.savez(out, params=T.params().astype(np.float32),
                         summary=json.dumps({"partial": True, "step": step, "heldout_loss": held}))
    theta = T.params()
    held = evaluate(T, cptr, ccol, hold_data, batch)
    # held-out solves: iterations with the trained network, coarse Jacobi and the exact inverse
    L.record(0)
    its = {"network": [], "jacobi": [], "exact": []}
    L.set_params(theta)
    for seed in seeds[-min(n_hold, 10):]:
        fam.assemble(seed, sigma)
        for mode, cm in (("network", rs.COARSE_NETWORK), ("jacobi", rs.COARSE_JACOBI), ("exact", rs.COARSE_EXACT)):
            L.set_cycle(1, 1, cm)
            its[mode].append(fam.solve(L).iterations)
    summary = {"holdout": n_hold, "pairs": int(sum(len(d[1]) for d in train_data)), "batch": batch,
               "steps": steps, "lr": lr, "heldout_loss_start": base, "heldout_loss_trained": held,
               "heldout_iterations": {k: float(np.mean(v)) for k, v in its.items()}, "curve": curve}
    say(f"held-out loss {base:.4f} -> {held:.4f}; iterations {summary['heldout_iterations']}")
    return theta, summary


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--sigma", type=float, default=2.5)
    ap.add_argument("--problems", type=int, default=2000)
    ap.add_argument("--seed0", type=int, default=10000)
    ap.add_argument("--record", type=int, default=64)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--init", default=None, help="start from these weights (.npz)")
    ap.add_argument("--rounds", type=int, default=1, help="training rounds; later rounds use on-policy pairs")
    args = ap.parse_args()
    rs = _package()
    rs.init(0)
    init = rs.load_weights(args.init)[0] if args.init else None
    train(rs, args.cells, args.sigma, args.problems, args.seed0, record=args.record, batch=args.batch,
          steps=args.steps, lr=args.lr, out=args.out, init=init, rounds=args.rounds)
    return 0


if __name__ == "__main__":
    sys.exit(main())
