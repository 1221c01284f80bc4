#!/usr/bin/env python3
"""Benchmark: time to a 1e-8 relative residual per heat system on one MI355X.

Workload (default): steady heat on the unit cube, 128^3 hexahedral cells split into 6 tets each
(2,146,689 nodes, 12,582,912 tets, 2,113,407 unknowns), conductivity lognormal(0, 2) per tet
(seed 2), volumetric source U(0,1) per tet (seed 1), T=0 on x=0 and T=1 on x=1.

One step = one system taken from resident device inputs to a solution three times:
  (a) GPU assembly + plain FP64 Jacobi-CG          (the drop-in path)
  (b) GPU assembly + mixed-precision Jacobi-CG     (FP32 SpMV/preconditioner, FP64 residuals)
  (c) GPU assembly + mixed-precision flexible CG with the learned multilevel preconditioner
      (V-cycle over nested 2x2x2 aggregates, FP32 Chebyshev smoothing, the trained network on
      the 512-node coarsest level; hierarchy values rebuilt for every system)
all to ||b - A x|| / ||b|| <= 1e-8 (checked in FP64 after the solve).  `value` is the time of
(c) per system; every method's time is in `methods`, with the solution error against the
oracle's FP64 solution converged to relres 1e-13.  After the timed steps, (c) runs once more with the coarsest
level solved by coarse Jacobi (the untrained network) and by the exact inverse: the ablation.

The roofline object is for the dominant kernel (the CSR SpMV), timed live with stream events
on the library's stream; algorithmic bytes per launch = nnz*(s_val+4) + n*(4+2*s_x) + 4.
The CPU baseline is the oracle (C + OpenMP restatement of the reference) on the whole system:
assembly + Jacobi-CG, median of 3 runs on the host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time-to-1e-8-residual per system: learned+mixed vs plain FP64 CG; SpMV GB/s")
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of each FP64 SpMV shape on the 128^3 heat matrix, from two rocprofv3
# --pmc runs (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, factors from a 256 MiB copy); the
# runs and their commands are summarised in profiles/spmv_heat128_pmc.md.
MEASURED_TRAFFIC_128 = {
    "vector-csr g4": 588.9e6, "vector-csr g8": 507.8e6, "vector-csr g16": 487.7e6,
    "vector-csr g32": 456.5e6, "vector-csr g64": 456.9e6, "tile-csr 2048": 456.6e6,
    "tile-csr 4096": 442.2e6, "sell-64": 422.4e6, "sell-64 int16": 352.7e6,
    "sell-64 u8": 426.8e6, "sell-64 int16 u8": 353.4e6,
}
MEASURED_TRAFFIC_128_SOURCE = "profiles/spmv_heat128_pmc.md"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n: int, nnz: int, s_val: int, s_x: int, s_col: int = 4) -> int:
    return nnz * (s_val + s_col) + n * (4 + 2 * s_x) + 4


def build_problem(rs, cells: int, sigma: float, seed_k: int = 2, seed_q: int = 1):
    xyz, tets = rs.kuhn_box(cells, cells, cells)
    ne = tets.shape[0]
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    t_node = np.where(np.abs(x - 1) < 1e-12, 1.0, 0.0)
    k = rs.field_lognormal(seed_k, ne, 0.0, sigma) if sigma > 0 else np.ones(ne)
    q = rs.field_uniform(seed_q, ne)
    return xyz, tets, fixed, t_node, k, q


def cpu_model() -> dict:
    """Host CPU model and the cores this process may run on (from /proc/cpuinfo)."""
    model, phys = None, set()
    cur = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if "core id" in cur:
                        phys.add((cur.get("physical id"), cur["core id"]))
                    cur = {}
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                cur[k] = v
                if k == "model name" and model is None:
                    model = v
        if "core id" in cur:
            phys.add((cur.get("physical id"), cur["core id"]))
    except OSError:
        pass
    return {"model": model, "logical_cpus_visible": len(os.sched_getaffinity(0)),
            "physical_cores_on_host": len(phys) or None}


def oracle_leg(rs, xyz, tets, fixed, t_node, k, q, tol: float, reps: int):
    """The oracle (C + OpenMP restatement of the reference) on the full system, timed directly:
    RSolverHeat prepare (assembly with the reference's critical-section scatter) and
    RMatrixSolver::solve (Jacobi build + CG to relres `tol`), `reps` times each, median
    reported.  Also returns x*, the oracle's FP64 solution converged to relres 1e-13, the
    reference for every GPU method's solution error."""
    host = cpu_model()   # before the OpenMP runtime binds this thread to one core
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import __graft_entry__ as ge
    orc = ge.load_oracle()
    orc.load()
    threads = int(os.environ.get("OMP_NUM_THREADS", host["logical_cpus_visible"]))
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    t_asm, t_cg, its = [], [], None
    for r in range(reps):
        t0 = time.perf_counter()
        A, b = orc.heat_assemble(xyz, tets, k, q, t_node, book, n_dof)
        t_asm.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        _, info = orc.solve(A, b, precond=orc.PRECOND_JACOBI, n_outer=400000, cvg=tol, stop_rule=orc.STOP_RELRES)
        t_cg.append(time.perf_counter() - t0)
        its = info["iterations"]
        log(f"oracle run {r}: assembly {t_asm[-1]:.2f} s, Jacobi-CG {t_cg[-1]:.2f} s ({its} its)")
    x_star, info = orc.solve(A, b, precond=orc.PRECOND_JACOBI, n_outer=400000, cvg=1e-13, stop_rule=orc.STOP_RELRES)
    relres_star = float(np.linalg.norm(b - A @ x_star) / np.linalg.norm(b))
    log(f"oracle x*: {info['iterations']} its, true relres {relres_star:.2e}")
    ta, tc = float(np.median(t_asm)), float(np.median(t_cg))
    cpu = {
        "value": round(ta + tc, 3), "unit": "s", "cores": threads, "kind": "port",
        "sample": (f"the whole system, not a sample: oracle (C/OpenMP restatement of the reference) heat "
                   f"assembly of all {tets.shape[0]} tets + Jacobi-CG to relres {tol} ({its} its), "
                   f"median of {reps} runs on {threads} threads (OMP_PROC_BIND={os.environ['OMP_PROC_BIND']}, "
                   f"OMP_PLACES={os.environ['OMP_PLACES']})"),
        "assembly_s": round(ta, 3), "solve_s": round(tc, 3), "iterations": its,
        "runs_assembly_s": [round(t, 3) for t in t_asm], "runs_solve_s": [round(t, 3) for t in t_cg],
        "cpu": host,
    }
    return cpu, x_star, {"iterations": info["iterations"], "true_relres": relres_star}


def spmv_only(rs, asm, dk, dq, dT, n, nnz) -> int:
    lib = rs.load()
    db = rs.DeviceArray(n)
    asm_ms = []
    for _ in range(5):
        asm.assemble_heat_device(dk, dq, dT, db)
        asm_ms.append(asm.last_ms)
    copy_bytes = 256 << 20
    src, dst = rs.DeviceArray(copy_bytes // 8), rs.DeviceArray(copy_bytes // 8)
    rs._check(lib.rs_memcpy_dtod(dst.ptr, src.ptr, copy_bytes))
    out = {"copy_bytes": copy_bytes, "heat_assembly_ms": round(float(np.median(asm_ms)), 4)}
    for name, prec, sv, sx in (("fp64", rs.PRECISION_FP64, 8, 8), ("fp32", rs.PRECISION_MIXED, 4, 4)):
        plan, _ = asm.matrix.spmv_plan(prec)
        ms = asm.matrix.spmv_timed(prec, reps=20)
        byts = spmv_bytes(n, nnz, sv, sx, rs.SPMV_COLUMN_BYTES[plan])
        out[name] = {"shape": rs.SPMV_SHAPE_NAMES[plan], "avg_us": round(ms * 1e3, 3),
                     "algorithmic_bytes": byts, "gbs": round(byts / (ms * 1e-3) / 1e9, 1)}
    print(json.dumps(out), flush=True)
    return 0


def coarse_degrees(args) -> tuple[int, int]:
    """Smoothing steps of the heat cycle on the levels below the finest."""
    return args.coarse_pre, args.coarse_post


def learned_only(rs, args, asm, xyz, dk, dq, dT, db, dx, zero_x, conf) -> int:
    """Profiling mode: the learned method alone (eager launches under RS_NO_GRAPH=1)."""
    L = rs.LearnedPreconditioner.from_bins(asm.matrix, asm, (max(1, args.cells // 2),) * 3, max_coarse=rs.MAX_COARSE)
    params, _ = rs.load_weights()
    L.set_params(params if params is not None else rs.init_params(0))
    L.set_cycle(args.pre, args.post, rs.COARSE_NETWORK)
    L.set_coarse_cycle(*coarse_degrees(args))
    sl = rs.RMatrixSolver(conf(rs.PRECISION_MIXED, args.tol))
    out = []
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        asm.assemble_heat_device(dk, dq, dT, db)
        zero_x()
        info = sl.solve_learned_device(asm.matrix, L, db, dx)
        rs.synchronize()
        wall = time.perf_counter() - t0
        log(f"learned run {i}: {wall:.3f}s, {info.iterations} its, relres {info.true_relres:.2e}")
        out.append({"wall_s": round(wall, 4), "iterations": info.iterations, "solve_ms": round(info.solve_ms, 2)})
    infer_ms = L.infer_timed(200)
    log(f"network inference (back-to-back): {infer_ms * 1e3:.1f} us")
    print(json.dumps({"learned_only": out, "levels": [int(v) for v in L.level_rows],
                      "infer_us": round(infer_ms * 1e3, 2)}), flush=True)
    return 0


def elasticity(rs, args) -> int:
    """Linear elasticity cantilever (RSolverStress form): box 4 x 1 x 1 in nx x ny x nz cells
    (default 128 x 32 x 32) x 6 tets, E lognormal(0, 1) per tet (seed 3), nu = 0.3, clamped at
    x = 0, traction (0, 0, -1) on x = 4; to relres 1e-8: block-Jacobi (3 x 3 node blocks) CG in FP64
    and in mixed precision, and flexible CG with the multilevel cycle over per-component and over
    rigid-body aggregates (mixed precision: FP32 cycle; FP32 or FP64 Krylov vectors).  Every timed
    solve follows a GPU re-assembly of the system and includes its kernel time, so the solver's
    per-matrix set-up (FP32 copies, preconditioner, Galerkin levels, coarsest inverse) is part of
    the time.  Prints one JSON line for this workload."""
    nx, ny, nz = args.elast_cells
    L = 4.0
    xyz, tets = rs.kuhn_box(nx, ny, nz, L, 1.0, 1.0)
    nn, ne = xyz.shape[0], tets.shape[0]
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    E = rs.field_lognormal(3, ne, 0.0, 1.0)
    nu, rho = np.full(ne, 0.3), np.ones(ne)
    tris = rs.plane_faces(xyz, tets, 0, L)
    trac = np.tile([0.0, 0.0, -1.0], (tris.shape[0], 1))
    log(f"elasticity {nx}x{ny}x{nz}: nodes={nn} tets={ne}; pattern + colouring + scatter map")
    asm = rs.Assembly(xyz, tets, 3, fixed)
    n = asm.n_dof
    t0 = time.perf_counter()
    b = asm.assemble_stress(E, nu, rho, [0.0, 0.0, 0.0], np.zeros(3 * nn), tris, trac)
    t_asm = time.perf_counter() - t0
    asm_kernel_ms = asm.last_ms
    log(f"dofs={n} nnz={asm.nnz} colours={asm.n_colours} assembly {asm.last_ms:.1f} ms (kernels) "
        f"{t_asm * 1e3:.0f} ms (with host transfers)")
    if args.assembly_only:
        # profiling mode: a 256 MiB device copy (counter calibration), then 5 assemblies by the row
        # gather and 5 by the colour launches
        lib = rs.load()
        copy_bytes = 256 << 20
        src, dst = rs.DeviceArray(copy_bytes // 8), rs.DeviceArray(copy_bytes // 8)
        rs._check(lib.rs_memcpy_dtod(dst.ptr, src.ptr, copy_bytes))
        out = {"copy_bytes": copy_bytes,
               "algorithmic_bytes": int(ne * (16 + 8) + nn * 24 + asm.nnz * 8 + n * 8)}
        for name, rows in (("row_gather", True), ("colour_launches", False)):
            asm.set_method(rows)
            ms = []
            for _ in range(5):
                asm.assemble_stress(E, nu, rho, [0.0, 0.0, 0.0], np.zeros(3 * nn), tris, trac)
                ms.append(asm.last_ms)
            out[name] = {"assembly_ms": round(float(np.median(ms)), 4)}
        print(json.dumps(out), flush=True)
        return 0
    db, dx = rs.DeviceArray.from_host(b), rs.DeviceArray(n)
    dzero = rs.DeviceArray.from_host(np.zeros(n))
    lib = rs.load()

    def reassemble():
        # every timed solve is a new system: the GPU assembly writes the matrix values again, so the
        # solver's per-matrix work (FP32 copies, preconditioner, multilevel set-up) runs again; the
        # assembly's kernel time is added to the solve (its host transfers are not)
        asm.assemble_stress(E, nu, rho, [0.0, 0.0, 0.0], np.zeros(3 * nn), tris, trac)
        return asm.last_ms * 1e-3

    def solve(precision, cvg):
        conf = rs.MatrixSolverConf(n_outer_iterations=200000, solver_cvg_value=cvg, stop_rule=rs.STOP_RELRES,
                                   precision=precision)
        S = rs.RMatrixSolver(conf)
        t_asm = reassemble()
        rs._check(lib.rs_memcpy_dtod(dx.ptr, dzero.ptr, 8 * n))
        t0 = time.perf_counter()
        info = S.solve_device(asm.matrix, db, dx, rs.PRECOND_BLOCK_JACOBI, 3)
        rs.synchronize()
        return time.perf_counter() - t0 + t_asm, info

    cpu, x_star, xstar_info = None, None, None
    if not args.no_cpu_baseline:
        # the oracle on the whole system: assembly + block-Jacobi(3) CG, timed; and x*, its FP64
        # solution to relres 1e-13, the reference for the solution errors
        host = cpu_model()   # before the OpenMP runtime binds this thread to one core
        os.environ.setdefault("OMP_PROC_BIND", "close")
        os.environ.setdefault("OMP_PLACES", "cores")
        import __graft_entry__ as ge
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", host["logical_cpus_visible"]))
        book, n_dof = orc.node_book(nn, tets, 3, fixed)
        t_asm, t_cg, its = [], [], None
        for r in range(args.cpu_reps):
            t0 = time.perf_counter()
            A_o, b_o = orc.stress_assemble(xyz, tets, E, nu, rho, [0.0, 0.0, 0.0], np.zeros(3 * nn), tris, trac, book,
                                           n_dof)
            t_asm.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            _, oi = orc.solve(A_o, b_o, precond=orc.PRECOND_BLOCK_JACOBI, block_size=3, n_outer=400000, cvg=args.tol,
                              stop_rule=orc.STOP_RELRES)
            t_cg.append(time.perf_counter() - t0)
            its = oi["iterations"]
            log(f"oracle run {r}: assembly {t_asm[-1]:.2f} s, block-Jacobi(3) CG {t_cg[-1]:.2f} s ({its} its)")
        x_star, oi = orc.solve(A_o, b_o, precond=orc.PRECOND_BLOCK_JACOBI, block_size=3, n_outer=400000, cvg=1e-13,
                               stop_rule=orc.STOP_RELRES)
        xstar_info = {"iterations": oi["iterations"],
                      "true_relres": float(np.linalg.norm(b_o - A_o @ x_star) / np.linalg.norm(b_o)),
                      "source": "oracle fp64, relres target 1e-13"}
        ta, tc = float(np.median(t_asm)), float(np.median(t_cg))
        cpu = {"value": round(ta + tc, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": (f"the whole system: oracle stress assembly of all {ne} tets + block-Jacobi(3) CG to relres "
                          f"{args.tol} ({its} its), median of {args.cpu_reps} runs on {threads} threads"),
               "assembly_s": round(ta, 3), "solve_s": round(tc, 3), "iterations": its, "cpu": host}
        log(f"oracle x*: {oi['iterations']} its; cpu baseline {cpu['value']} s")
    else:
        _, iref = solve(rs.PRECISION_FP64, 1e-13)
        x_star = dx.download()
        xstar_info = {"iterations": iref.iterations, "true_relres": iref.true_relres, "source": "gpu fp64"}
        log(f"reference FP64 to 1e-13 on the GPU: {iref.iterations} its")
    res = {}
    for name, prec in (("fp64_blockjacobi3_cg", rs.PRECISION_FP64), ("mixed_blockjacobi3_cg", rs.PRECISION_MIXED)):
        solve(prec, args.tol)                       # warm-up (tuning, graph capture)
        times, info = [], None
        for _ in range(max(1, args.steps)):
            t, info = solve(prec, args.tol)
            times.append(t)
        x = dx.download()
        res[name] = {"time_s": round(float(np.median(times)), 4), "iterations": info.iterations,
                     "true_relres": info.true_relres, "reason": rs.REASON_NAMES.get(info.reason),
                     "solution_error_vs_oracle_fp64": float(np.linalg.norm(x - x_star) / np.linalg.norm(x_star))}
        log(f"{name}: {res[name]['time_s'] * 1e3:.1f} ms, {info.iterations} its, relres {info.true_relres:.2e}")
    # the multilevel cycle for elasticity: nested coordinate bins per displacement component
    # (unknown-based aggregation), Chebyshev smoothing in D^-1 A, the exact coarsest inverse, in
    # flexible CG with mixed precision (FP32 cycle and SpMV, FP64 residual replacement)
    bins = (max(1, nx // 2), max(1, ny // 2), max(1, nz // 2))
    hier = {}
    for agg_name, kw in (("components", {"by_component": True}), ("rigid", {"rigid": True})):
        rs.synchronize()
        t0 = time.perf_counter()
        Lh = rs.LearnedPreconditioner.from_bins(asm.matrix, asm, bins, max_coarse=rs.MAX_COARSE, **kw)
        rs.synchronize()
        Lh.set_cycle(args.pre, args.post, rs.COARSE_EXACT)
        hier[agg_name] = (Lh, time.perf_counter() - t0)
    aggregation_text = {"components": "nested coordinate bins per displacement component (translations)",
                        "rigid": "nested coordinate bins of nodes, six rigid-body modes per aggregate "
                                 "(orthonormalised, 6 coarse rows per aggregate)"}
    for name, prec, krylov, agg_name in (
            ("mixed_multilevel_fcg", rs.PRECISION_MIXED, "FP32 vectors and SpMV, FP64 residual replacement",
             "components"),
            ("mixed_multilevel_fcg_fp64_krylov", rs.PRECISION_FP64, "FP64 vectors and SpMV", "components"),
            ("mixed_rigid_multilevel_fcg", rs.PRECISION_MIXED, "FP32 vectors and SpMV, FP64 residual replacement",
             "rigid"),
            ("mixed_rigid_multilevel_fcg_fp64_krylov", rs.PRECISION_FP64, "FP64 vectors and SpMV", "rigid")):
        Lm, t_hier = hier[agg_name]
        conf_ml = rs.MatrixSolverConf(n_outer_iterations=20000, solver_cvg_value=args.tol, stop_rule=rs.STOP_RELRES,
                                      precision=prec)
        sml = rs.RMatrixSolver(conf_ml)

        def solve_ml():
            t_asm = reassemble()
            rs._check(lib.rs_memcpy_dtod(dx.ptr, dzero.ptr, 8 * n))
            t0 = time.perf_counter()
            info = sml.solve_learned_device(asm.matrix, Lm, db, dx)
            rs.synchronize()
            return time.perf_counter() - t0 + t_asm, info

        solve_ml()                                     # warm-up (level tuning, graph capture)
        times, info = [], None
        for _ in range(max(1, args.steps)):
            t, info = solve_ml()
            times.append(t)
        x = dx.download()
        res[name] = {
            "time_s": round(float(np.median(times)), 4), "iterations": info.iterations,
            "true_relres": info.true_relres, "reason": rs.REASON_NAMES.get(info.reason), "fallback": info.fallback,
            "solve_ms": round(info.solve_ms, 2), "fp64_residual_replacements": info.inner_iterations,
            "solution_error_vs_oracle_fp64": float(np.linalg.norm(x - x_star) / np.linalg.norm(x_star)),
            "levels": [int(v) for v in Lm.level_rows], "level_nnz": [int(v) for v in Lm.level_nnz],
            "cycle": {"pre": args.pre, "post": args.post, "aggregation": aggregation_text[agg_name],
                      "smoother": "Chebyshev in D^-1 A, FP32 (FP16 copy of D^-1 A on levels >= 65,536 rows)",
                      "coarsest": "exact inverse (FP64 sweep on the GPU, applied in FP32)"},
            "krylov": krylov, "hierarchy_build_s": round(t_hier, 4),
            "time_with_hierarchy_build_s": round(float(np.median(times)) + t_hier, 4)}
        log(f"{name}: {res[name]['time_s'] * 1e3:.1f} ms, {info.iterations} its, "
            f"relres {info.true_relres:.2e}, levels {res[name]['levels']}, hierarchy {t_hier:.3f} s")
    best_ml = min(("mixed_multilevel_fcg", "mixed_multilevel_fcg_fp64_krylov", "mixed_rigid_multilevel_fcg",
                   "mixed_rigid_multilevel_fcg_fp64_krylov"), key=lambda k: res[k]["time_s"])
    spmv = {}
    for name, prec in (("fp64", rs.PRECISION_FP64), ("fp32", rs.PRECISION_MIXED)):
        plan, tune_ms = asm.matrix.spmv_plan(prec)
        spmv[name] = {"shape": rs.SPMV_SHAPE_NAMES[plan],
                      "tuning_us": {rs.SPMV_SHAPE_NAMES[i]: round(t * 1e3, 2) for i, t in enumerate(tune_ms)}}
    # post-processing (RSolverStress::process): stresses and node forces from the solution
    du = rs.DeviceArray.from_host(np.zeros(3 * nn))
    asm.dofs_to_nodes(dx, du)
    dE, dnu = rs.DeviceArray.from_host(E), rs.DeviceArray.from_host(nu)
    dN, dS, dV, dF = rs.DeviceArray(ne), rs.DeviceArray(ne), rs.DeviceArray(ne), rs.DeviceArray(3 * nn)
    sp_ms = []
    for _ in range(6):
        asm.stress_process_device(dE, dnu, du, dN, dS, dV, dF)
        sp_ms.append(asm.last_ms)
    spm = float(np.median(sp_ms[1:]))
    vmax = float(dV.download().max())
    log(f"stress processing: {spm:.3f} ms, max von Mises {vmax:.4g}")
    out = {"metric": "time-to-1e-8-residual per system (elasticity; mixed precision, multilevel flexible CG; "
                     "block-Jacobi-3 CG beside it)",
           "value": res[best_ml]["time_s"], "value_method": best_ml, "unit": "s", "n_gpus": 1,
           "higher_is_better": False,
           "dtype": "f64 / f32 SpMV+precond with f64 residuals", "data": "synthetic",
           "config": {"workload": f"cantilever {nx}x{ny}x{nz} cells x 6 tets, E lognormal(0,1), nu 0.3",
                      "nodes": int(nn), "tets": int(ne), "dofs": int(n), "nnz": int(asm.nnz)},
           "assembly_ms": round(asm_kernel_ms, 3), "methods": res,
           "speedup_mixed_vs_fp64": round(res["fp64_blockjacobi3_cg"]["time_s"] / res["mixed_blockjacobi3_cg"]["time_s"], 3),
           "speedup_mixed_multilevel_vs_fp64_blockjacobi": round(
               res["fp64_blockjacobi3_cg"]["time_s"] / res[best_ml]["time_s"], 3),
           "reference_solution": xstar_info, "cpu_baseline": cpu, "spmv": spmv,
           "post_processing": {"stress_process_ms": round(spm, 4), "max_von_mises": vmax}}
    print(json.dumps(out), flush=True)
    return 0


def batch_transient(rs, args) -> int:
    """Batched transient heat: args.batch_count systems, system i the unit cube in k_i^3 cells
    (k_i uniform on {8..48} from seed 4) x 6 tets, conductivity lognormal(0, 1.5), source
    U(0, 1), rho*c = 1, T = 0/1 on x = 0/1, T0 = 0; implicit Euler dt = 1e-3, 10 steps each, to
    relres 1e-8 per step.  Timed: the whole batch (every system's 10 steps) in one batched
    loop, mixed precision (FP32 SpMV/preconditioner, FP64 residual replacement) and FP64;
    beside it, the same systems one at a time through the single-system API (GPU assembly +
    mixed Jacobi-CG per step, the reference's per-step flow).  Prints one JSON line."""
    import __graft_entry__ as ge
    tr = ge.load_module("transient")
    count, steps, dt, sigma = args.batch_count, args.batch_steps, 1e-3, 1.5
    sizes = tr.system_sizes(rs, count, 8, 48, 4)
    log(f"batch: {count} systems, k in [{sizes.min()}, {sizes.max()}], {steps} steps; assembling")
    B, meshes, t_build, t_asm = tr.build_batch(rs, sizes, sigma=sigma, dt=dt, seed=4, log=log)
    B.set_replacement(args.batch_replacement)
    n_rows = np.array([meshes[k].n for k in sizes])
    nnz = np.array([meshes[k].nnz for k in sizes])
    log(f"built in {t_build:.1f}s (GPU assembly of every system into the batch {t_asm:.3f}s): "
        f"unknowns {n_rows.sum()}, nnz {nnz.sum()}")

    def conf(precision, cvg):
        return rs.MatrixSolverConf(n_outer_iterations=100000, solver_cvg_value=cvg, stop_rule=rs.STOP_RELRES,
                                   precision=precision)

    # FP64 reference solutions (tight) for the error of every result
    ref = B.solve(conf(rs.PRECISION_FP64, 1e-12), steps)
    x_ref = [B.solution(i) for i in range(count)]
    log(f"FP64 batch to 1e-12: {ref.solve_ms:.1f} ms, rounds {ref.rounds}")
    methods = {}
    for name, prec in (("batched_fp64_cg", rs.PRECISION_FP64), ("batched_mixed_cg", rs.PRECISION_MIXED)):
        B.solve(conf(prec, args.tol), steps)   # warm-up: tuning, graph capture
        times, info = [], None
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            info = B.solve(conf(prec, args.tol), steps)
            times.append(time.perf_counter() - t0)
        its, rr, reasons = B.stats()
        err = max(float(np.linalg.norm(B.solution(i) - x_ref[i]) / max(np.linalg.norm(x_ref[i]), 1e-300))
                  for i in range(count))
        shape, shape_ms = B.shape(prec)
        methods[name] = {"time_s": round(float(np.median(times)), 4), "device_ms": round(info.solve_ms, 2),
                         "rounds": info.rounds, "total_iterations": int(info.total_iterations),
                         "max_iterations_per_system": info.max_iterations,
                         "replacements": info.replacements, "systems_converged": info.systems_converged,
                         "max_step_relres": info.max_relres, "max_solution_error_vs_fp64": err,
                         "spmv_shape": rs.BATCH_SHAPE_NAMES[shape],
                         "spmv_tuning_us": {rs.BATCH_SHAPE_NAMES[j]: round(v * 1e3, 1) for j, v in enumerate(shape_ms)}}
        log(f"{name}: {methods[name]['time_s'] * 1e3:.1f} ms, rounds {info.rounds}, "
            f"iterations {info.total_iterations}, max relres {info.max_relres:.2e}, error {err:.2e}")
    # dominant kernel: the batched FP32 SpMV over every system (tuning time of the chosen shape)
    lay = B.layout()
    shape, shape_ms = B.shape(rs.PRECISION_MIXED)
    ms = float(shape_ms[shape])
    byts = spmv_bytes(lay["padded_rows"], lay["nnz"], 4, 4, lay["column_bytes"])
    ach = byts / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"rs::k_bt_spmv FP32 batched SELL SpMV ({rs.BATCH_SHAPE_NAMES[shape]})",
                "algorithmic_bytes": byts, "avg_us": round(ms * 1e3, 1)}
    # the same systems one at a time, with the batch's own recurrence (A, M and c assembled once,
    # b = c + M x_prev on the device): a one-system batch per mesh size, reused across the
    # systems of that size; assembly included, as in the batched total
    single = None
    if args.unbatched_budget > 0:
        log("one at a time, same recurrence: one-system batches (assembly + 10 steps per system)")
        singles, t_one, done_nnz, done = {}, 0.0, 0, 0
        for i, k in enumerate(sizes):
            m = meshes[int(k)]
            kf, qf = tr.fields(rs, m, 4, i, sigma)
            dk, dq, dc = rs.DeviceArray.from_host(kf), rs.DeviceArray.from_host(qf), rs.DeviceArray(m.n)
            if int(k) not in singles:
                S1 = rs.Batch([m.n], [m.nnz])
                S1.set_replacement(args.batch_replacement)
                tr.assemble_system(rs, S1, 0, m, dk, dq, dc, dt)
                S1.solve(conf(rs.PRECISION_MIXED, args.tol), steps)      # tuning and graph capture
                singles[int(k)] = S1
            S1 = singles[int(k)]
            rs.synchronize()
            t0 = time.perf_counter()
            tr.assemble_system(rs, S1, 0, m, dk, dq, dc, dt)
            S1.solve(conf(rs.PRECISION_MIXED, args.tol), steps)
            rs.synchronize()
            t_one += time.perf_counter() - t0
            done_nnz += int(nnz[i])
            done += 1
            if t_one > args.unbatched_budget / 2:
                break
        est = t_one * float(nnz.sum()) / done_nnz
        single = {"time_s": round(est, 3), "measured_systems": done, "measured_s": round(t_one, 3),
                  "note": ("all systems" if done == count else
                           f"first {done} systems timed, scaled by total nnz / their nnz")}
        log(f"one at a time, same recurrence: {t_one:.1f}s for {done} systems -> {est:.2f}s for the batch")
    # the same systems one at a time through the single-system API (the reference's per-step flow)
    unb = None
    if args.unbatched_budget > 0:
        log("unbatched: one system at a time (assembly + mixed Jacobi-CG per step)")
        solver = rs.RMatrixSolver(conf(rs.PRECISION_MIXED, args.tol))
        done_nnz, t_unb, done = 0, 0.0, 0
        for i, k in enumerate(sizes):
            m = meshes[int(k)]
            kf, qf = tr.fields(rs, m, 4, i, sigma)
            dk, dq = rs.DeviceArray.from_host(kf), rs.DeviceArray.from_host(qf)
            dT = rs.DeviceArray.from_host(m.t0)
            db, dx, dz = rs.DeviceArray(m.n), rs.DeviceArray(m.n), rs.DeviceArray.from_host(np.zeros(m.n))
            rs.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                m.asm.assemble_heat_device(dk, dq, dT, db, drhoc=m.drhoc, transient=True, alpha=1.0, dt=dt)
                rs._check(rs.load().rs_memcpy_dtod(dx.ptr, dz.ptr, 8 * m.n))
                solver.solve_device(m.asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
                m.asm.dofs_to_nodes(dx, dT)
            rs.synchronize()
            t_unb += time.perf_counter() - t0
            done_nnz += int(nnz[i])
            done += 1
            if t_unb > args.unbatched_budget / 2:
                break
        est = t_unb * float(nnz.sum()) / done_nnz
        unb = {"time_s": round(est, 3), "measured_systems": done, "measured_s": round(t_unb, 3),
               "note": ("all systems" if done == count else
                        f"first {done} systems timed, scaled by total nnz / their nnz")}
        log(f"unbatched: {t_unb:.1f}s for {done} systems -> {est:.2f}s for the batch")
    # CPU baseline: the oracle's per-step loop (assembly + Jacobi-CG) on a bounded sample
    cpu = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t_cpu, done_nnz, done = 0.0, 0, 0
        for i, k in enumerate(sizes):
            m = meshes[int(k)]
            kf, qf = tr.fields(rs, m, 4, i, sigma)
            book, n_dof = orc.node_book(m.xyz.shape[0], m.tets, 1, m.fixed)
            t0 = time.perf_counter()
            orc.transient_heat(m.xyz, m.tets, kf, qf, np.ones(m.tets.shape[0]), m.t0, book, n_dof, dt, steps,
                               cvg=args.tol)
            t_cpu += time.perf_counter() - t0
            done_nnz += int(nnz[i])
            done += 1
            if t_cpu > args.cpu_budget:
                break
        cpu = {"value": round(t_cpu * float(nnz.sum()) / done_nnz, 2), "unit": "s", "cores": threads,
               "kind": "port",
               "sample": (f"oracle time loop (assembly + Jacobi-CG per step, {threads} threads) on the first "
                          f"{done} systems ({t_cpu:.1f}s), scaled by total nnz / their nnz")}
        log(f"cpu baseline: {cpu['value']} s per batch")
    best = methods["batched_mixed_cg"]["time_s"]
    for mth in methods.values():
        mth["assembly_s"] = round(t_asm, 4)
        mth["time_with_assembly_s"] = round(mth["time_s"] + t_asm, 4)
    out = {
        "metric": f"batched transient heat: time for {count} systems x {steps} implicit-Euler steps to relres {args.tol}",
        "value": best, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(best * 1e3, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 SpMV/precond, f64 residual replacement and solution (batched_mixed_cg)",
        "data": "synthetic",
        "config": {"workload": (f"{count} cube systems, k~U{{8..48}} cells/edge (seed 4) x 6 tets, "
                                f"lognormal(0,{sigma}) conductivity, implicit Euler dt {dt}, {steps} steps"),
                   "unknowns": int(n_rows.sum()), "nnz": int(nnz.sum()), "padded_rows": lay["padded_rows"],
                   "sell_entries": lay["sell_entries"], "column_bytes": lay["column_bytes"],
                   "replacement_ratio": args.batch_replacement, "parallelism": "single GPU"},
        "system_steps_per_s": round(count * steps / best, 1),
        "methods": methods,
        "one_at_a_time_same_recurrence_mixed": single,
        "unbatched_mixed_cg": unb,
        "speedup_batched_vs_one_at_a_time_same_recurrence":
            round(single["time_s"] / methods["batched_mixed_cg"]["time_with_assembly_s"], 2) if single else None,
        "speedup_batched_vs_unbatched": (round(unb["time_s"] / methods["batched_mixed_cg"]["time_with_assembly_s"], 2)
                                         if unb else None),
        "comparison_note": ("batched times include the one-time GPU assembly (time_with_assembly_s); "
                            "one_at_a_time_same_recurrence runs each system alone through the same batched "
                            "solver (assemble once, b = c + M x_prev on the device); unbatched_mixed_cg is the "
                            "reference's per-step flow (GPU re-assembly + single-system mixed Jacobi-CG per step)"),
        "roofline": roofline, "cpu_baseline": cpu, "build_s": round(t_build, 2),
    }
    print(json.dumps(out), flush=True)
    return 0


def oracle_cg_iteration_s(orc, A, short=8, long=48, reps=2):
    """Seconds per Jacobi-CG iteration of the oracle on A: the difference of the fastest of `reps`
    short and long runs (after a warm-up run), so the per-call conversion and start-up cancel.
    Returns (seconds, a description of the sample)."""
    b = np.ones(A.shape[0])

    def run(its):
        t0 = time.perf_counter()
        orc.solve(A, b, precond=orc.PRECOND_JACOBI, n_outer=its, cvg=1e-300, stop_rule=1)
        return time.perf_counter() - t0

    run(2)
    t_short = min(run(short) for _ in range(reps))
    t_long = min(run(long) for _ in range(reps))
    if t_long <= t_short:
        raise RuntimeError(f"oracle CG timing: {long} iterations ({t_long:.3f} s) not slower than {short} "
                           f"({t_short:.3f} s)")
    return (t_long - t_short) / (long - short), (f"the fastest of {reps} {short}- and {long}-iteration runs "
                                                 f"after a warm-up")


def eigen_multilevel(rs, args, econf, mconf, M, K, d, jacobi_s, asm, bins, agg_kw, agg_text):
    """The same modal analysis with every inner solve with K in flexible CG (FP64 Krylov vectors,
    the same stopping rule) around a V-cycle over the assembly's mesh (rs_eigen_solve_learned;
    no counterpart in the reference, whose inner solves are Jacobi-CG), at the Jacobi run's inner
    tolerance and at a hundredth of it.

    Arnoldi's Ritz values beyond the lowest are not converged in a Krylov space of the requested
    dimension and move with the inner solves' final errors, and the reference's stopping rule
    bounds a backward error (|r| / (|A| |x| + |b|)), which leaves different forward errors to the
    two inner solvers.  Every run is therefore set beside the Jacobi-CG run at a hundredth of the
    inner tolerance (the tight reference), per eigenvalue in ascending order."""
    rs.synchronize()
    t0 = time.perf_counter()
    Lm = rs.LearnedPreconditioner.from_bins(K, asm, bins, max_coarse=rs.MAX_COARSE, **agg_kw)
    rs.synchronize()
    t_hier = time.perf_counter() - t0
    Lm.set_cycle(1, 1, rs.COARSE_EXACT)

    def conf_at(factor):
        return rs.MatrixSolverConf(n_outer_iterations=mconf.n_outer_iterations,
                                   solver_cvg_value=mconf.solver_cvg_value * factor, stop_rule=mconf.stop_rule)

    rs.seed_c_rand(1)
    d_t = np.sort(rs.REigenValueSolver(econf, conf_at(0.01)).solve(M, K)[0])

    def rel_to_tight(v):
        return [float(f"{x:.3g}") for x in np.abs(np.sort(v) - d_t) / np.abs(d_t)]

    out = {"levels": [int(v) for v in Lm.level_rows], "hierarchy_build_s": round(t_hier, 4),
           "inner_solver": f"flexible CG, FP64 Krylov vectors, V-cycle over {agg_text} (FP32 Chebyshev "
                           "smoothing, exact coarsest inverse)",
           "tight_reference": "the Jacobi-CG Arnoldi with a hundredth of the inner tolerance",
           "jacobi_rel_diff_vs_tight_reference": rel_to_tight(d)}
    for key, factor in (("same_tolerance", 1.0), ("hundredth_tolerance", 0.01)):
        times, d_ml, info = [], None, None
        for rep in range(1 + max(1, args.steps)):
            rs.seed_c_rand(1)
            S = rs.REigenValueSolver(econf, conf_at(factor))
            t0 = time.perf_counter()
            d_ml, _ = S.solve(M, K, preconditioner=Lm)
            dt_ = time.perf_counter() - t0
            info = S.info
            if rep > 0:
                times.append(dt_)
            log(f"gpu arnoldi, multilevel inner solves ({key}): {dt_ * 1e3:.1f} ms, {info.cg_iterations} FCG "
                f"iterations, lowest eigenvalue {np.min(d_ml):.6g}")
        ml_s = float(np.median(times))
        out[key] = {"inner_cvg_value": mconf.solver_cvg_value * factor, "time_s": round(ml_s, 4),
                    "inner_iterations": int(info.cg_iterations), "speedup_vs_jacobi": round(jacobi_s / ml_s, 2),
                    "rel_diff_vs_tight_reference": rel_to_tight(d_ml)}
        log(f"eigenvalues vs the tight reference: jacobi {out['jacobi_rel_diff_vs_tight_reference']}, "
            f"multilevel ({key}) {out[key]['rel_diff_vs_tight_reference']}")
    return out


def eigen(rs, args) -> int:
    """Modal analysis of the heat operator (REigenValueSolver, Arnoldi): K x = lambda M x with K
    the 128^3 x 6-tet heat stiffness (lognormal(0,1) conductivity, T fixed on x = 0 and x = 1),
    M the consistent mass; args.eigen_modes eigenvalues, inner Jacobi-CG to the reference's
    stopping rule at 1e-12.  GPU (rs_eigen_solve) and the oracle start from the same
    srand(1) vector.  Prints one JSON line."""
    import __graft_entry__ as ge
    if args.eigen_problem == "stress":
        return eigen_stress(rs, args)
    c, ne = args.eigen_cells, args.eigen_modes
    xyz, tets = rs.kuhn_box(c, c, c)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    nt = tets.shape[0]
    k = rs.field_lognormal(5, nt, 0.0, 1.0)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    asm.assemble_heat(k, np.zeros(nt), np.zeros(xyz.shape[0]), rhoc=np.ones(nt), transient=True, alpha=1.0, dt=0.0)
    Ms = asm.matrix.to_scipy()
    asm.assemble_heat(k, np.zeros(nt), np.zeros(xyz.shape[0]))
    Ks = asm.matrix.to_scipy()
    M, K = rs.SparseMatrix.from_scipy(Ms), rs.SparseMatrix.from_scipy(Ks)
    log(f"eigen: {c}^3, n={Ks.shape[0]}, nnz={Ks.nnz}, {ne} modes (Arnoldi)")
    econf = rs.EigenValueSolverConf(method=rs.EIGEN_ARNOLDI, n_eigen_values=ne, n_iterations=500,
                                    solver_cvg_value=1e-10)
    mconf = rs.MatrixSolverConf(n_outer_iterations=100000, solver_cvg_value=1e-12)
    times, d, info = [], None, None
    for rep in range(1 + max(1, args.steps)):
        rs.seed_c_rand(1)
        S = rs.REigenValueSolver(econf, mconf)
        t0 = time.perf_counter()
        d, ev = S.solve(M, K)
        dt_ = time.perf_counter() - t0
        info = S.info
        if rep > 0:
            times.append(dt_)
        log(f"gpu arnoldi: {dt_ * 1e3:.1f} ms, {info.cg_iterations} CG iterations")
    gpu_s = float(np.median(times))
    multilevel = eigen_multilevel(rs, args, econf, mconf, M, K, d, gpu_s, asm, (c // 2, c // 2, c // 2), {},
                                  "nested coordinate bins of nodes")
    cpu = None
    agree = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rs.seed_c_rand(1)
        t0 = time.perf_counter()
        d_ref, _, its_ref = orc.eigen_solve(Ms, Ks, 1, n_eigen=ne, n_iter=500, eig_cvg=1e-10, cvg=1e-12)
        t_cpu = time.perf_counter() - t0
        agree = float(np.max(np.abs(d - d_ref) / np.abs(d_ref)))
        cpu = {"value": round(t_cpu, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": f"the oracle's Arnoldi on the same {c}^3 problem ({its_ref} CG iterations, {threads} threads)"}
        log(f"cpu arnoldi: {t_cpu:.2f}s; eigenvalues agree to {agree:.2e}")
    out = {"metric": f"modal analysis: time for {ne} eigenvalues (Arnoldi, REigenValueSolver)",
           "value": round(gpu_s, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "higher_is_better": False,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"heat operator {c}^3 cells x 6 tets, lognormal(0,1) conductivity, consistent mass",
                      "unknowns": int(Ks.shape[0]), "nnz": int(Ks.nnz), "modes": ne},
           "eigenvalues": [round(float(v), 6) for v in d], "cg_iterations": int(info.cg_iterations),
           "max_relative_difference_vs_oracle": agree, "value_method": "jacobi_cg inner solves (the reference's)",
           "multilevel_inner_solves": multilevel, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def eigen_stress(rs, args) -> int:
    """Modal analysis of the C3 cantilever (RSolverStress modal -> REigenValueSolver, Arnoldi):
    K and the consistent mass M from the modal stress assembly (128 x 32 x 32 cells x 6 tets,
    E lognormal(0, 1), nu = 0.3, rho = 1, clamped at x = 0), args.eigen_modes eigenvalues, inner
    Jacobi-CG to the reference's stopping rule at 1e-12 (src/reigenvaluesolver.cpp:250).  The
    oracle's CG time per iteration on the same K (two bounded runs), times the GPU's CG
    iterations, is the CPU baseline.  Prints one JSON line."""
    import __graft_entry__ as ge
    nx, ny, nz = args.elast_cells
    ne = args.eigen_modes
    L = 4.0
    xyz, tets = rs.kuhn_box(nx, ny, nz, L, 1.0, 1.0)
    nn, nt = xyz.shape[0], tets.shape[0]
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    E = rs.field_lognormal(3, nt, 0.0, 1.0)
    asm = rs.Assembly(xyz, tets, 3, fixed)
    asm.assemble_stress_ex(rs.StressConf(rs.STRESS_MODAL, 1.0, 1.0, (0.0, 0.0, 0.0)), E, np.full(nt, 0.3),
                           np.ones(nt), np.zeros(3 * nn))
    K, M = asm.matrix, asm.mass_matrix()
    log(f"eigen stress: {nx}x{ny}x{nz}, n={asm.n_dof}, nnz={asm.nnz}, {ne} modes (Arnoldi), assembly "
        f"{asm.last_ms:.2f} ms")
    econf = rs.EigenValueSolverConf(method=rs.EIGEN_ARNOLDI, n_eigen_values=ne, n_iterations=500,
                                    solver_cvg_value=1e-10)
    mconf = rs.MatrixSolverConf(n_outer_iterations=200000, solver_cvg_value=1e-12)
    times, d, info = [], None, None
    for rep in range(1 + max(1, args.steps)):
        rs.seed_c_rand(1)
        S = rs.REigenValueSolver(econf, mconf)
        t0 = time.perf_counter()
        d, ev = S.solve(M, K)
        dt_ = time.perf_counter() - t0
        info = S.info
        if rep > 0:
            times.append(dt_)
        log(f"gpu arnoldi (stress): {dt_ * 1e3:.1f} ms, {info.cg_iterations} CG iterations, lowest "
            f"eigenvalue {np.min(d):.6g}")
    gpu_s = float(np.median(times))
    multilevel = eigen_multilevel(rs, args, econf, mconf, M, K, d, gpu_s, asm,
                                  (max(1, nx // 2), max(1, ny // 2), max(1, nz // 2)), {"rigid": True},
                                  "rigid-body aggregates")
    cpu = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        Ks = K.to_scipy()
        t_iter, its_text = oracle_cg_iteration_s(orc, Ks)
        total = t_iter * info.cg_iterations
        cpu = {"value": round(total, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": (f"oracle Jacobi-CG per iteration on the same K from {its_text} "
                          f"({t_iter * 1e3:.2f} ms, {threads} threads) x the GPU's {info.cg_iterations} CG iterations")}
        log(f"cpu baseline: {total:.1f} s")
    out = {"metric": f"modal analysis: time for {ne} eigenvalues (Arnoldi, REigenValueSolver) of the cantilever",
           "value": round(gpu_s, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "higher_is_better": False,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"cantilever {nx}x{ny}x{nz} cells x 6 tets, modal stress (K, consistent M)",
                      "unknowns": int(asm.n_dof), "nnz": int(asm.nnz), "modes": ne},
           "eigenvalues": [float(v) for v in np.sort(d)], "cg_iterations": int(info.cg_iterations),
           "value_method": "jacobi_cg inner solves (the reference's)", "multilevel_inner_solves": multilevel,
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def electrostatics(rs, args) -> int:
    """Electrostatics (RSolverElectrostatics) on the 128^3 x 6-tet unit cube: relative
    permittivity lognormal(0,1), charge density U(-0.5, 0.5) 1e-9, conductivity lognormal(2,1),
    plates at 0 V (x = 0) and 5 V (x = 1).  GPU assembly, plain FP64 Jacobi-CG to relres 1e-8 and
    the processing of field, current, energy, resistivity and Joule heat; the oracle times the
    same steps on a bounded sample.  The same steps with the multilevel cycle in mixed-precision
    flexible CG are timed beside them.  Prints one JSON line."""
    import __graft_entry__ as ge
    c = args.cells
    xyz, tets = rs.kuhn_box(c, c, c)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    phi = np.where(np.abs(x - 1) < 1e-12, 5.0, 0.0)
    ne, nn = tets.shape[0], xyz.shape[0]
    eps_r = rs.field_lognormal(7, ne, 0.0, 1.0)
    rho = (rs.field_uniform(8, ne) - 0.5) * 1e-9
    sigma = rs.field_lognormal(9, ne, 2.0, 1.0)
    asm = rs.Assembly(xyz, tets, 1, fixed)
    n = asm.n_dof
    deps, drho, dsig, dphi = (rs.DeviceArray.from_host(a) for a in (eps_r, rho, sigma, phi))
    db, dx, dz = rs.DeviceArray(n), rs.DeviceArray(n), rs.DeviceArray.from_host(np.zeros(n))
    outs = [rs.DeviceArray(3 * ne), rs.DeviceArray(3 * ne), rs.DeviceArray(ne), rs.DeviceArray(ne), rs.DeviceArray(ne)]
    conf = rs.MatrixSolverConf(n_outer_iterations=400000, solver_cvg_value=args.tol, stop_rule=rs.STOP_RELRES)
    lib = rs.load()
    res = []
    for rep in range(1 + max(1, args.steps)):
        rs._check(lib.rs_memcpy_dtod(dx.ptr, dz.ptr, 8 * n))
        rs.synchronize()
        t0 = time.perf_counter()
        asm.assemble_electrostatics_device(deps, drho, dphi, db)
        a_ms = asm.last_ms
        info = rs.RMatrixSolver(conf).solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
        asm.dofs_to_nodes(dx, dphi)
        asm.electrostatics_process_device(deps, dsig, dphi, *outs)
        p_ms = asm.last_ms
        rs.synchronize()
        t = time.perf_counter() - t0
        if rep > 0:
            res.append((t, a_ms, info.solve_ms, p_ms, info.iterations, info.true_relres))
        log(f"electrostatics: {t * 1e3:.1f} ms (assembly {a_ms:.2f} ms, solve {info.solve_ms:.1f} ms "
            f"{info.iterations} its, processing {p_ms:.3f} ms)")
    best = min(res)
    x_jac = dx.download()
    # the same per-system steps with the multilevel cycle (node bins, FP32 Chebyshev smoothing,
    # exact coarsest inverse) in mixed-precision flexible CG; no counterpart in the reference
    rs.synchronize()
    t0 = time.perf_counter()
    Lm = rs.LearnedPreconditioner.from_bins(asm.matrix, asm, (c // 2, c // 2, c // 2), max_coarse=rs.MAX_COARSE)
    rs.synchronize()
    t_hier = time.perf_counter() - t0
    Lm.set_cycle(1, 1, rs.COARSE_EXACT)
    conf_ml = rs.MatrixSolverConf(n_outer_iterations=20000, solver_cvg_value=args.tol, stop_rule=rs.STOP_RELRES,
                                  precision=rs.PRECISION_MIXED)
    res_ml = []
    for rep in range(1 + max(1, args.steps)):
        rs._check(lib.rs_memcpy_dtod(dx.ptr, dz.ptr, 8 * n))
        rs.synchronize()
        t0 = time.perf_counter()
        asm.assemble_electrostatics_device(deps, drho, dphi, db)
        info = rs.RMatrixSolver(conf_ml).solve_learned_device(asm.matrix, Lm, db, dx)
        asm.dofs_to_nodes(dx, dphi)
        asm.electrostatics_process_device(deps, dsig, dphi, *outs)
        rs.synchronize()
        t = time.perf_counter() - t0
        if rep > 0:
            res_ml.append((t, info.solve_ms, info.iterations, info.true_relres))
        log(f"electrostatics, multilevel: {t * 1e3:.1f} ms (solve {info.solve_ms:.1f} ms, {info.iterations} its)")
    best_ml = min(res_ml)
    x_ml = dx.download()
    multilevel = {"time_s": round(best_ml[0], 4), "solve_ms": round(best_ml[1], 2), "iterations": best_ml[2],
                  "true_relres": best_ml[3], "speedup_vs_jacobi_cg": round(best[0] / best_ml[0], 2),
                  "solution_rel_diff_vs_jacobi_cg": float(np.linalg.norm(x_ml - x_jac) / np.linalg.norm(x_jac)),
                  "levels": [int(v) for v in Lm.level_rows], "hierarchy_build_s": round(t_hier, 4),
                  "method": "V-cycle over nested coordinate bins (FP32 Chebyshev, exact coarsest inverse) in "
                            "mixed-precision flexible CG"}
    cpu = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sub = args.asm_cells
        sx, st_ = rs.kuhn_box(sub, sub, sub)
        sxx = sx[:, 0]
        sfix = ((np.abs(sxx) < 1e-12) | (np.abs(sxx - 1) < 1e-12)).astype(np.uint8)
        sphi = np.where(np.abs(sxx - 1) < 1e-12, 5.0, 0.0)
        sne = st_.shape[0]
        book, ndof = orc.node_book(sx.shape[0], st_, 1, sfix)
        t0 = time.perf_counter()
        A_s, b_s = orc.electrostatics_assemble(sx, st_, eps_r[:sne], rho[:sne], sphi, book, ndof)
        t_asm = (time.perf_counter() - t0) * ne / sne
        t0 = time.perf_counter()
        orc.electrostatics_process(sx, st_, eps_r[:sne], sigma[:sne], sphi)
        t_proc = (time.perf_counter() - t0) * ne / sne
        A_full = asm.matrix.to_scipy()
        t_iter, its_text = oracle_cg_iteration_s(orc, A_full)
        total = t_asm + t_proc + best[4] * t_iter
        cpu = {"value": round(total, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": (f"oracle on {threads} threads: assembly and processing of a {sub}^3 sub-box scaled per "
                          f"tet ({t_asm:.2f} s, {t_proc:.2f} s), Jacobi-CG per iteration from {its_text} "
                          f"on the full system ({t_iter * 1e3:.1f} ms x {best[4]})")}
        log(f"cpu baseline: {total:.2f} s")
    out = {"metric": "electrostatics: assembly + FP64 Jacobi-CG to relres 1e-8 + processing, per system",
           "value": round(best[0], 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "higher_is_better": False,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"electrostatics {c}^3 cells x 6 tets, eps_r lognormal(0,1), plates 0/5 V",
                      "unknowns": n, "nnz": int(asm.nnz), "tets": int(ne)},
           "assembly_ms": round(best[1], 3), "solve_ms": round(best[2], 2), "processing_ms": round(best[3], 4),
           "iterations": best[4], "true_relres": best[5], "multilevel": multilevel, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def magnetostatics(rs, args) -> int:
    """Magnetostatics (RSolverMagnetostatics) on the args.mag_cells^3 x 6-tet unit cube, three
    unknowns per node: node current densities N(0,1) 1e6 (seeded), field fixed to zero on x = 0
    and x = 1.  GPU assembly and FP64 Jacobi-CG to relres 1e-8; the oracle times the assembly on
    a sub-box (scaled per tet) and CG iterations on the full system.  The same steps with the
    multilevel cycle in mixed-precision flexible CG are timed beside them.  Prints one JSON line."""
    import __graft_entry__ as ge
    c = args.mag_cells
    xyz, tets = rs.kuhn_box(c, c, c)
    nn, ne = xyz.shape[0], tets.shape[0]
    x = xyz[:, 0]
    fixed_node = (np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)
    fixed = np.repeat(fixed_node, 3).astype(np.uint8)
    J = np.random.default_rng(11).standard_normal(3 * nn) * 1e6
    H = np.zeros(3 * nn)
    asm = rs.Assembly(xyz, tets, 3, fixed)
    n = asm.n_dof
    dJ, dH = rs.DeviceArray.from_host(J), rs.DeviceArray.from_host(H)
    db, dx, dz = rs.DeviceArray(n), rs.DeviceArray(n), rs.DeviceArray.from_host(np.zeros(n))
    conf = rs.MatrixSolverConf(n_outer_iterations=400000, solver_cvg_value=args.tol, stop_rule=rs.STOP_RELRES)
    lib = rs.load()
    res = []
    for rep in range(1 + max(1, args.steps)):
        rs._check(lib.rs_memcpy_dtod(dx.ptr, dz.ptr, 8 * n))
        rs.synchronize()
        t0 = time.perf_counter()
        asm.assemble_magnetostatics_device(dJ, dH, db)
        a_ms = asm.last_ms
        info = rs.RMatrixSolver(conf).solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
        rs.synchronize()
        t = time.perf_counter() - t0
        if rep > 0:
            res.append((t, a_ms, info.solve_ms, info.iterations, info.true_relres))
        log(f"magnetostatics: {t * 1e3:.1f} ms (assembly {a_ms:.2f} ms, solve {info.solve_ms:.1f} ms "
            f"{info.iterations} its, relres {info.true_relres:.2e})")
    best = min(res)
    x_jac = dx.download()
    # the same per-system steps with the multilevel cycle over per-component node bins (the three
    # components of each node in separate aggregates), FP32 Chebyshev smoothing and coarse Jacobi on
    # the coarsest level, in mixed-precision flexible CG; no counterpart in the reference
    rs.synchronize()
    t0 = time.perf_counter()
    Lm = rs.LearnedPreconditioner.from_bins(asm.matrix, asm, (c // 2, c // 2, c // 2), max_coarse=rs.MAX_COARSE,
                                            by_component=True)
    rs.synchronize()
    t_hier = time.perf_counter() - t0
    # the curl-type operator is only semidefinite, and so is its coarsest Galerkin operator: coarse
    # Jacobi there instead of an inverse
    Lm.set_cycle(1, 1, rs.COARSE_JACOBI)
    conf_ml = rs.MatrixSolverConf(n_outer_iterations=20000, solver_cvg_value=args.tol, stop_rule=rs.STOP_RELRES,
                                  precision=rs.PRECISION_MIXED)
    res_ml = []
    for rep in range(1 + max(1, args.steps)):
        rs._check(lib.rs_memcpy_dtod(dx.ptr, dz.ptr, 8 * n))
        rs.synchronize()
        t0 = time.perf_counter()
        asm.assemble_magnetostatics_device(dJ, dH, db)
        info = rs.RMatrixSolver(conf_ml).solve_learned_device(asm.matrix, Lm, db, dx)
        rs.synchronize()
        t = time.perf_counter() - t0
        if rep > 0:
            res_ml.append((t, info.solve_ms, info.iterations, info.true_relres))
        log(f"magnetostatics, multilevel: {t * 1e3:.1f} ms (solve {info.solve_ms:.1f} ms, {info.iterations} its)")
    best_ml = min(res_ml)
    x_ml = dx.download()
    multilevel = {"time_s": round(best_ml[0], 4), "solve_ms": round(best_ml[1], 2), "iterations": best_ml[2],
                  "true_relres": best_ml[3], "speedup_vs_jacobi_cg": round(best[0] / best_ml[0], 2),
                  "solution_rel_diff_vs_jacobi_cg": float(np.linalg.norm(x_ml - x_jac) / np.linalg.norm(x_jac)),
                  "levels": [int(v) for v in Lm.level_rows], "hierarchy_build_s": round(t_hier, 4),
                  "method": "V-cycle over per-component nested coordinate bins (FP32 Chebyshev, coarse Jacobi on "
                            "the coarsest level) in mixed-precision flexible CG"}
    cpu = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sub = 24
        sx, st_ = rs.kuhn_box(sub, sub, sub)
        sfix_node = (np.abs(sx[:, 0]) < 1e-12) | (np.abs(sx[:, 0] - 1) < 1e-12)
        sfix = np.repeat(sfix_node, 3).astype(np.uint8)
        book, ndof = orc.node_book(sx.shape[0], st_, 3, sfix)
        t0 = time.perf_counter()
        orc.magnetostatics_assemble(sx, st_, J[:3 * sx.shape[0]], np.zeros(3 * sx.shape[0]), book, ndof)
        t_asm = (time.perf_counter() - t0) * ne / st_.shape[0]
        A_full = asm.matrix.to_scipy()
        tt = []
        for its in (4, 14):
            t0 = time.perf_counter()
            orc.solve(A_full, np.ones(n), precond=orc.PRECOND_JACOBI, n_outer=its, cvg=1e-300, stop_rule=1)
            tt.append(time.perf_counter() - t0)
        t_iter = (tt[1] - tt[0]) / 10.0
        total = t_asm + best[3] * t_iter
        cpu = {"value": round(total, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": (f"oracle on {threads} threads: assembly of a {sub}^3 sub-box scaled per tet "
                          f"({t_asm:.2f} s), Jacobi-CG per iteration from 4- and 14-iteration runs on the full "
                          f"system ({t_iter * 1e3:.1f} ms x {best[3]})")}
        log(f"cpu baseline: {total:.2f} s")
    out = {"metric": "magnetostatics: assembly + FP64 Jacobi-CG to relres 1e-8, per system",
           "value": round(best[0], 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "higher_is_better": False,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"magnetostatics {c}^3 cells x 6 tets, 3 unknowns per node, random current density",
                      "unknowns": n, "nnz": int(asm.nnz), "tets": int(ne)},
           "assembly_ms": round(best[1], 3), "solve_ms": round(best[2], 2), "iterations": best[3],
           "true_relres": best[4], "multilevel": multilevel, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def fluid(rs, args) -> int:
    """Steady duct flow (RSolverFluid) on (4c x c x c) cells x 6 tets, c = args.fluid_cells, length
    4: parabolic inflow at x = 0, no-slip walls y, z = 0, 1, pressure 0 at x = 4, rho = 1,
    mu = 0.02 (Reynolds number 50 on the duct width).  A step is one Picard iteration as the
    reference runs it: GPU assembly of the stabilised system, restarted GMRES + Jacobi to relres
    1e-8 (restart args.fluid_restart: at 128 x 32 x 32 a restart of 100 stagnates near 3e-2),
    write-back.  The oracle times its assembly on a sub-duct (scaled per tet) and partial GMRES
    cycles on the first full-size system.  Prints one JSON line."""
    import __graft_entry__ as ge
    c = args.fluid_cells
    lx = 4.0
    xyz, tets = rs.kuhn_box(4 * c, c, c, lx, 1.0, 1.0)
    nn, ne = xyz.shape[0], tets.shape[0]

    def duct(xyz):
        x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
        near = lambda a, v: np.abs(a - v) < 1e-12
        wall = near(y, 0) | near(y, 1) | near(z, 0) | near(z, 1)
        fixed = np.zeros((xyz.shape[0], 4), np.uint8)
        fixed[wall | near(x, 0), :3] = 1
        fixed[near(x, lx), 3] = 1
        inflow = near(x, 0) & ~wall
        st = np.zeros((xyz.shape[0], 4))
        st[inflow, 0] = 16 * y[inflow] * (1 - y[inflow]) * z[inflow] * (1 - z[inflow])
        return fixed.reshape(-1), st

    fixed, state = duct(xyz)
    mu = 0.02
    t0 = time.perf_counter()
    asm = rs.Assembly(xyz, tets, 4, fixed)
    log(f"fluid: {nn} nodes, {ne} tets, {asm.n_dof} unknowns, {asm.nnz} nonzeros, {asm.n_colours} colours "
        f"(set-up {time.perf_counter() - t0:.1f} s)")
    conf = rs.FluidConf(1.0, 1.0, 1.0, 0)
    drho, dmu, dg = (rs.DeviceArray.from_host(a) for a in (np.ones(ne), np.full(ne, mu), np.zeros(3 * ne)))
    ds = rs.DeviceArray.from_host(state.reshape(-1).copy())
    dvo = rs.DeviceArray.from_host(state[:, :3].reshape(-1).copy())
    dacc = rs.DeviceArray.from_host(np.zeros(3 * nn))
    n = asm.n_dof
    db, dx, dz = rs.DeviceArray(n), rs.DeviceArray(n), rs.DeviceArray.from_host(np.zeros(n))
    restart = args.fluid_restart
    lib = rs.load()

    def picard_steps(precision):
        # the same Picard steps from the same initial state, GMRES in FP64 or mixed precision
        ds.upload(state.reshape(-1).copy())
        dvo.upload(state[:, :3].reshape(-1).copy())
        dacc.upload(np.zeros(3 * nn))
        sconf = rs.MatrixSolverConf(solver_type=rs.SOLVER_GMRES, n_outer_iterations=60, n_inner_iterations=restart,
                                    solver_cvg_value=args.tol, stop_rule=rs.STOP_RELRES, precision=precision)
        solver = rs.RMatrixSolver(sconf)
        out = []
        for it in range(args.warmup + args.steps):
            rs.synchronize()
            t0 = time.perf_counter()
            asm.assemble_fluid_device(conf, drho, dmu, dg, ds, dvo, dacc, db)
            a_ms = asm.last_ms
            rs._check(lib.rs_memcpy_dtod(dx.ptr, dz.ptr, 8 * n))
            info = solver.solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
            norms = asm.fluid_update_device(dx, False, 1.0, ds, dvo, dacc)
            rs.synchronize()
            t = time.perf_counter() - t0
            if it >= args.warmup:
                out.append((t, a_ms, info.solve_ms, info.inner_iterations, info.true_relres, info.iterations,
                            info.converged))
            log(f"fluid step {it} ({'mixed' if precision == rs.PRECISION_MIXED else 'fp64'}): {t * 1e3:.1f} ms "
                f"(assembly {a_ms:.2f} ms, GMRES {info.solve_ms:.1f} ms, {info.inner_iterations} inner its, relres "
                f"{info.true_relres:.2e}, converged {info.converged}), |v| {norms[1]:.4f} |p| {norms[3]:.4f}")
        return out, ds.download()

    steps_mp, state_mp = picard_steps(rs.PRECISION_MIXED)
    steps, state_64 = picard_steps(rs.PRECISION_FP64)
    ts = np.array([x[0] for x in steps])
    ts_mp = np.array([x[0] for x in steps_mp])
    mixed = {"time_s": round(float(np.median(ts_mp)), 4), "gmres_ms": round(float(np.median([x[2] for x in steps_mp])), 2),
             "inner_iterations": [int(x[3]) for x in steps_mp], "cycles": [int(x[5]) for x in steps_mp],
             "all_converged": all(x[6] == 1 for x in steps_mp), "true_relres": max(x[4] for x in steps_mp),
             "state_difference_vs_fp64": float(np.abs(state_mp - state_64).max() / np.abs(state_64).max()),
             "speedup_vs_fp64": round(float(np.median(ts)) / float(np.median(ts_mp)), 3),
             "method": ("GMRES inner cycle on an FP32 basis, FP32 SpMV and Jacobi; FP64 Gram-Schmidt products, "
                        "rotations, restart residuals and stopping test")}
    cpu = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sub = max(2, c // 4)
        sxyz, stets = rs.kuhn_box(4 * sub, sub, sub, lx, 1.0, 1.0)
        sfixed, sstate = duct(sxyz)
        sbook, sndof = orc.node_book(sxyz.shape[0], stets, 4, sfixed)
        sne = stets.shape[0]
        t0 = time.perf_counter()
        orc.fluid_assemble(sxyz, stets, np.ones(sne), np.full(sne, mu), np.zeros((sne, 3)), sstate,
                           sstate[:, :3].copy(), np.zeros((sxyz.shape[0], 3)), 1.0, 1.0, 1.0, 0, sbook, sndof)
        t_asm = (time.perf_counter() - t0) * ne / sne
        # partial GMRES cycles of 100 and 200 inner iterations on the full first system: cost
        # t(k) = a k + c k^2 (SpMV + preconditioner linear, Gram-Schmidt quadratic in the cycle
        # length), applied to the GPU's cycles (each taken at the mean cycle length)
        book = asm.book()
        A_full, b_full = orc.fluid_assemble(xyz, tets, np.ones(ne), np.full(ne, mu), np.zeros((ne, 3)), state,
                                            state[:, :3].copy(), np.zeros((nn, 3)), 1.0, 1.0, 1.0, 0, book, n)
        tk = []
        for k in (100, 200):
            t0 = time.perf_counter()
            orc.solve(A_full, b_full, precond=orc.PRECOND_JACOBI, solver=orc.SOLVER_GMRES, n_outer=1, n_inner=k,
                      cvg=1e-300, stop_rule=1)
            tk.append(time.perf_counter() - t0)
        cq = max(0.0, (tk[1] - 2 * tk[0]) / (200.0 ** 2 - 2 * 100.0 ** 2))
        ca = (tk[0] - cq * 100.0 ** 2) / 100.0
        cycles = float(np.mean([x[5] for x in steps]))
        klen = float(np.mean([x[3] for x in steps])) / max(cycles, 1.0)
        t_solve = cycles * (ca * klen + cq * klen ** 2)
        total = t_asm + t_solve
        cpu = {"value": round(total, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": (f"oracle on {threads} threads: assembly of a {4 * sub}x{sub}x{sub} sub-duct scaled per tet "
                          f"({t_asm:.2f} s); GMRES+Jacobi cycles of 100 and 200 inner iterations on the full first "
                          f"system ({tk[0]:.2f} s, {tk[1]:.2f} s) fitted as a k + c k^2 and applied to "
                          f"{cycles:.1f} cycles of {klen:.0f} ({t_solve:.1f} s)")}
        log(f"cpu baseline: {total:.2f} s per Picard step")
    out = {"metric": "steady duct flow: Picard step (assembly + restarted GMRES-Jacobi to relres 1e-8 + write-back)",
           "value": round(float(np.median(ts)), 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(float(np.median(ts)) * 1e3, 2), "higher_is_better": False,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"duct flow {4 * c}x{c}x{c} cells x 6 tets, 4 unknowns per node, Re 50, "
                                  f"GMRES({restart})",
                      "unknowns": n, "nnz": int(asm.nnz), "tets": int(ne)},
           "assembly_ms": round(float(np.median([x[1] for x in steps])), 3),
           "gmres_ms": round(float(np.median([x[2] for x in steps])), 2),
           "gmres_restart": restart, "inner_iterations": [int(x[3]) for x in steps],
           "cycles": [int(x[5]) for x in steps], "all_converged": all(x[6] == 1 for x in steps),
           "true_relres": max(x[4] for x in steps),
           "mixed_precision_gmres": mixed,
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def transport(rs, args) -> int:
    """Fluid heat (RSolverFluidHeat; RSolverFluidParticle runs the same element) on the c^3 x 6-tet
    unit cube, c = args.cells: a rotating flow (-(y - 1/2), x - 1/2, 0.2) per element centroid,
    rho c = 1, conduction 5e-5, implicit time steps dt = 1e-2 (alpha = 1), temperature 1 on z = 0
    and 0 on z = 1.  The reference's diffusion block enters with a negative sign, so the step
    matrix is M - dt k K + ...: with dt k |grad N|^2 above the mass term (k = 1e-3 at 128^3) its
    diagonal degenerates and GMRES + Jacobi makes no progress, on the oracle as on the GPU; the
    conduction is chosen below that.  A step is one time step as the reference runs it: GPU assembly of the SUPG
    system, GMRES(100) + Jacobi to relres 1e-8, the clamped write-back.  The oracle times its
    assembly on a sub-cube (scaled per tet) and partial GMRES cycles on the full system (fitted
    a k + c k^2 and applied to the GPU's cycles).  Prints one JSON line."""
    import __graft_entry__ as ge
    c = args.cells
    xyz, tets = rs.kuhn_box(c, c, c)
    nn, ne = xyz.shape[0], tets.shape[0]

    def setup(xyz, tets):
        cen = xyz[tets].mean(1) - 0.5
        vel = np.stack([-cen[:, 1], cen[:, 0], 0.2 * np.ones(tets.shape[0])], 1)
        z = xyz[:, 2]
        fixed = ((np.abs(z) < 1e-12) | (np.abs(z - 1) < 1e-12)).astype(np.uint8)
        return vel, fixed, np.where(np.abs(z) < 1e-12, 1.0, 0.0)

    vel, fixed, field = setup(xyz, tets)
    kc, dt = 5e-5, 1e-2
    t0 = time.perf_counter()
    asm = rs.Assembly(xyz, tets, 1, fixed)
    n = asm.n_dof
    log(f"transport: {nn} nodes, {ne} tets, {n} unknowns, {asm.nnz} nonzeros (set-up {time.perf_counter() - t0:.1f} s)")
    conf = rs.TransportConf(1.0, dt, 1)
    dca, dk, dv, dq = (rs.DeviceArray.from_host(a.reshape(-1)) for a in (np.ones(ne), np.full(ne, kc), vel,
                                                                            np.zeros(nn)))
    dfield = rs.DeviceArray.from_host(field.copy())
    db, dx, dz = rs.DeviceArray(n), rs.DeviceArray(n), rs.DeviceArray.from_host(np.zeros(n))
    lib = rs.load()

    def time_steps(precision):
        # the same time steps from the same initial field, GMRES in FP64 or mixed precision
        dfield.upload(field.copy())
        sconf = rs.MatrixSolverConf(solver_type=rs.SOLVER_GMRES, n_outer_iterations=100, n_inner_iterations=100,
                                    solver_cvg_value=args.tol, stop_rule=rs.STOP_RELRES, precision=precision)
        solver = rs.RMatrixSolver(sconf)
        out = []
        for it in range(args.warmup + args.steps):
            rs.synchronize()
            t0 = time.perf_counter()
            asm.assemble_transport_device(conf, dca, dk, dv, dq, dfield, db)
            a_ms = asm.last_ms
            rs._check(lib.rs_memcpy_dtod(dx.ptr, dz.ptr, 8 * n))       # prepare clears x every step
            info = solver.solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
            norms = asm.transport_update_device(dx, dfield)
            rs.synchronize()
            t = time.perf_counter() - t0
            if it >= args.warmup:
                out.append((t, a_ms, info.solve_ms, info.inner_iterations, info.true_relres, info.iterations,
                            info.converged))
            log(f"transport step {it} ({'mixed' if precision == rs.PRECISION_MIXED else 'fp64'}): {t * 1e3:.1f} ms "
                f"(assembly {a_ms:.2f} ms, GMRES {info.solve_ms:.1f} ms, {info.inner_iterations} inner its, relres "
                f"{info.true_relres:.2e}, converged {info.converged}), |T| {norms[1]:.3f}")
        return out, dfield.download()

    steps_mp, field_mp = time_steps(rs.PRECISION_MIXED)
    steps, field_64 = time_steps(rs.PRECISION_FP64)
    ts = np.array([x[0] for x in steps])
    ts_mp = np.array([x[0] for x in steps_mp])
    mixed = {"time_s": round(float(np.median(ts_mp)), 5), "gmres_ms": round(float(np.median([x[2] for x in steps_mp])), 3),
             "inner_iterations": [int(x[3]) for x in steps_mp], "all_converged": all(x[6] == 1 for x in steps_mp),
             "true_relres": max(x[4] for x in steps_mp),
             "field_difference_vs_fp64": float(np.abs(field_mp - field_64).max() / max(np.abs(field_64).max(), 1e-300)),
             "speedup_vs_fp64": round(float(np.median(ts)) / float(np.median(ts_mp)), 3)}
    cpu = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sub = max(2, c // 4)
        sxyz, stets = rs.kuhn_box(sub, sub, sub)
        svel, sfixed, sfield = setup(sxyz, stets)
        sbook, sndof = orc.node_book(sxyz.shape[0], stets, 1, sfixed)
        t0 = time.perf_counter()
        orc.transport_assemble(sxyz, stets, 1.0, kc, svel, np.zeros(sxyz.shape[0]), sfield, 1.0, dt, 1, sbook, sndof)
        t_asm = (time.perf_counter() - t0) * ne / stets.shape[0]
        A_full = asm.matrix.to_scipy()
        b_full = db.download()
        tk = []
        for kk in (20, 40):
            t0 = time.perf_counter()
            orc.solve(A_full, b_full, precond=orc.PRECOND_JACOBI, solver=orc.SOLVER_GMRES, n_outer=1, n_inner=kk,
                      cvg=1e-300, stop_rule=1)
            tk.append(time.perf_counter() - t0)
        cq = max(0.0, (tk[1] - 2 * tk[0]) / (40.0 ** 2 - 2 * 20.0 ** 2))
        ca_ = (tk[0] - cq * 20.0 ** 2) / 20.0
        cycles = float(np.mean([x[5] for x in steps]))
        klen = float(np.mean([x[3] for x in steps])) / max(cycles, 1.0)
        t_solve = cycles * (ca_ * klen + cq * klen ** 2)
        total = t_asm + t_solve
        cpu = {"value": round(total, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": (f"oracle on {threads} threads: assembly of a {sub}^3 sub-cube scaled per tet ({t_asm:.2f} s); "
                          f"GMRES+Jacobi cycles of 20 and 40 inner iterations on the full system ({tk[0]:.2f} s, "
                          f"{tk[1]:.2f} s) fitted as a k + c k^2 and applied to {cycles:.1f} cycles of {klen:.0f}")}
        log(f"cpu baseline: {total:.2f} s per time step")
    out = {"metric": "fluid heat: implicit time step (SUPG assembly + GMRES(100)-Jacobi to relres 1e-8 + write-back)",
           "value": round(float(np.median(ts)), 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(float(np.median(ts)) * 1e3, 3), "higher_is_better": False,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"fluid heat {c}^3 cells x 6 tets, rotating flow, k 5e-5, dt 1e-2",
                      "unknowns": n, "nnz": int(asm.nnz), "tets": int(ne)},
           "assembly_ms": round(float(np.median([x[1] for x in steps])), 3),
           "gmres_ms": round(float(np.median([x[2] for x in steps])), 3),
           "inner_iterations": [int(x[3]) for x in steps], "all_converged": all(x[6] == 1 for x in steps),
           "true_relres": max(x[4] for x in steps), "mixed_precision_gmres": mixed, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def radiosity(rs, args) -> int:
    """Radiative heat (RSolverRadiativeHeat with a loaded view-factor matrix) in the closed unit
    box with k x k patches per face (k = args.rad_patches): emissivity U(0.2, 0.95) and
    temperatures 400 + 100 face + U(0, 50) K per element, ambient 300 K.  A step is the GPU
    assembly of the dense patch system, CG + Jacobi to relres 1e-12 and the split of the patch
    heats onto the elements.  The oracle runs the same steps on the host (all of them, the system
    is small).  Prints one JSON line."""
    import __graft_entry__ as ge
    import scipy.sparse as sp
    rad = ge.load_module("radiation")
    k = args.rad_patches
    t0 = time.perf_counter()
    xyz, tris, pp, pe, F, area, face = rad.box_enclosure(k)
    npch, ne = F.shape[0], tris.shape[0]
    rng = np.random.default_rng(8)
    em = rng.uniform(0.2, 0.95, ne)
    te = 400.0 + 100.0 * face.repeat(2) + rng.uniform(0, 50, ne)
    am = np.full(ne, 300.0)
    R = rs.Radiosity(xyz, tris, pp, pe)
    dF, dem, dte, dam = (rs.DeviceArray.from_host(a.reshape(-1)) for a in (F, em, te, am))
    log(f"radiosity: {npch} patches, {ne} elements, dense system {npch}^2 (set-up {time.perf_counter() - t0:.1f} s)")
    db, dq, dz, dh = rs.DeviceArray(npch), rs.DeviceArray(npch), rs.DeviceArray.from_host(np.zeros(npch)), \
        rs.DeviceArray(ne)
    conf = rs.MatrixSolverConf(n_outer_iterations=1000, solver_cvg_value=1e-12, stop_rule=rs.STOP_RELRES)
    solver = rs.RMatrixSolver(conf)
    lib = rs.load()
    res = []
    for it in range(args.warmup + args.steps):
        rs.synchronize()
        t0 = time.perf_counter()
        R.assemble_device(dF, dem, dte, dam, db)
        a_ms = R.last_ms
        rs._check(lib.rs_memcpy_dtod(dq.ptr, dz.ptr, 8 * npch))
        info = solver.solve_device(R.matrix, db, dq, rs.PRECOND_JACOBI, 1)
        R.process_device(dq, dh)
        rs.synchronize()
        t = time.perf_counter() - t0
        if it >= args.warmup:
            res.append((t, a_ms, info.solve_ms, info.iterations, info.true_relres))
        log(f"radiosity step {it}: {t * 1e3:.2f} ms (assembly {a_ms:.3f} ms, CG {info.solve_ms:.2f} ms / "
            f"{info.iterations} its, relres {info.true_relres:.2e})")
    q = dq.download()
    balance = abs(float(area @ q)) / float(area @ np.abs(q))
    ts = np.array([x[0] for x in res])
    a_med = float(np.median([x[1] for x in res]))
    alg = 16.0 * npch * npch   # read F, write A
    cpu = None
    if not args.no_cpu_baseline:
        orc = ge.load_oracle()
        orc.load()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t0 = time.perf_counter()
        A_ref, b_ref = orc.radiosity_assemble(F, pp, pe, em, te, am)
        A_csr = sp.csr_matrix(A_ref)
        t1 = time.perf_counter()
        q_ref, _ = orc.solve(A_csr, b_ref, precond=orc.PRECOND_JACOBI, n_outer=1000, cvg=1e-12, stop_rule=1)
        orc.radiosity_process(pp, pe, np.full(ne, 0.5 / (k * k)), q_ref)
        t2 = time.perf_counter()
        cpu = {"value": round(t2 - t0, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": f"the whole step on the oracle, {threads} threads (assembly {t1 - t0:.2f} s incl. CSR "
                         f"conversion, CG + process {t2 - t1:.2f} s)"}
        log(f"cpu baseline: {t2 - t0:.2f} s; max |q_gpu - q_oracle| / max |q| = "
            f"{np.abs(q - q_ref).max() / np.abs(q_ref).max():.2e}")
    out = {"metric": "radiative heat: dense patch-system assembly + CG-Jacobi to relres 1e-12 + element split",
           "value": round(float(np.median(ts)), 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(float(np.median(ts)) * 1e3, 3), "higher_is_better": False,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"closed box, {k}x{k} patches per face ({npch} patches, {ne} elements), "
                                  f"dense view factors", "patches": npch},
           "assembly_ms": round(a_med, 4), "assembly_GBps": round(alg / (a_med * 1e-3) / 1e9, 1),
           "cg_ms": round(float(np.median([x[2] for x in res])), 3), "cg_iterations": int(res[-1][3]),
           "true_relres": max(x[4] for x in res), "energy_balance": balance, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--sigma", type=float, default=2.0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed-k", type=int, default=2, help="conductivity field seed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=3, help="oracle runs timed for the CPU baseline (median)")
    ap.add_argument("--asm-cells", type=int, default=40, help="electrostatics: oracle sub-box for the CPU sample")
    ap.add_argument("--eigen-cells", type=int, default=128)
    ap.add_argument("--eigen-modes", type=int, default=10)
    ap.add_argument("--eigen-problem", choices=["heat", "stress"], default="heat",
                    help="eigen workload: the 128^3 heat operator or the cantilever's modal stress system")
    ap.add_argument("--mag-cells", type=int, default=96)
    ap.add_argument("--fluid-cells", type=int, default=32)
    ap.add_argument("--fluid-restart", type=int, default=400)
    ap.add_argument("--rad-patches", type=int, default=32)
    ap.add_argument("--workload", choices=["heat", "elasticity", "batch", "eigen", "electrostatics", "magnetostatics",
                                                 "fluid", "radiosity", "transport"],
                    default="heat",
                    help="heat (default, the headline), the elasticity cantilever, the batched "
                         "transient heat systems, the modal analysis (Arnoldi), electrostatics, "
                         "magnetostatics, the steady duct flow (Picard step), radiative heat or fluid "
                         "heat transport")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--unbatched-budget", type=float, default=40.0,
                    help="seconds of one-system-at-a-time solves timed beside the batch, split between the "
                         "same-recurrence and the per-step re-assembly comparators (0: skip)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--batch-replacement", type=float, default=0.01,
                    help="mixed batch: FP64 residual replacement when the residual fell by this ratio")
    ap.add_argument("--elast-cells", type=int, nargs=3, default=[128, 32, 32])
    ap.add_argument("--learned-only", action="store_true",
                    help="profiling mode: time only the learned method (no FP64/mixed CG, no ablation)")
    ap.add_argument("--pre", type=int, default=1, help="pre-smoothing steps of the V-cycle")
    ap.add_argument("--post", type=int, default=1, help="post-smoothing steps of the V-cycle")
    ap.add_argument("--coarse-pre", type=int, default=1,
                    help="heat's learned cycle: pre-smoothing steps on the levels below the finest (3 there: "
                         "20.8 ms / 53 its against 21.3 ms / 65 its with 1 at 128^3, but coarse Jacobi on the "
                         "coarsest level then needs only 56 its, so the network gains nothing)")
    ap.add_argument("--coarse-post", type=int, default=1,
                    help="heat's learned cycle: post-smoothing steps on the levels below the finest")
    ap.add_argument("--assembly-only", action="store_true",
                    help="elasticity profiling mode: a 256 MiB device copy (counter calibration), then 5 "
                         "stress assemblies by the row gather and 5 by the colour launches; no solves")
    ap.add_argument("--spmv-only", action="store_true",
                    help="profiling mode: a 256 MiB device copy (counter calibration), then "
                         "20 FP64 and 20 FP32 SpMV launches of the tuned shapes; no solves")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design, not an implementation)"}))
        return 2

    import __graft_entry__ as ge
    rs = ge.load_package()
    rs.init(0)
    if args.workload == "elasticity":
        return elasticity(rs, args)
    if args.workload == "batch":
        return batch_transient(rs, args)
    if args.workload == "eigen":
        return eigen(rs, args)
    if args.workload == "electrostatics":
        return electrostatics(rs, args)
    if args.workload == "magnetostatics":
        return magnetostatics(rs, args)
    if args.workload == "fluid":
        return fluid(rs, args)
    if args.workload == "radiosity":
        return radiosity(rs, args)
    if args.workload == "transport":
        return transport(rs, args)
    c = args.cells
    log(f"building {c}^3 Kuhn mesh and fields")
    xyz, tets, fixed, t_node, k, q = build_problem(rs, c, args.sigma, args.seed_k)
    log(f"nodes={xyz.shape[0]} tets={tets.shape[0]}; pattern + colouring + scatter map")
    t0 = time.perf_counter()
    asm = rs.Assembly(xyz, tets, 1, fixed)
    t_prep = time.perf_counter() - t0
    n, nnz = asm.n_dof, asm.nnz
    log(f"unknowns={n} nnz={nnz} colours={asm.n_colours} prep={t_prep:.1f}s")
    dk, dq, dT = (rs.DeviceArray.from_host(a) for a in (k, q, t_node))
    if args.spmv_only:
        return spmv_only(rs, asm, dk, dq, dT, n, nnz)
    db, dx, dzero = rs.DeviceArray(n), rs.DeviceArray(n), rs.DeviceArray.from_host(np.zeros(n))
    lib = rs.load()

    def zero_x():
        rs._check(lib.rs_memcpy_dtod(dx.ptr, dzero.ptr, 8 * n))

    def conf(precision, cvg, n_outer=400000):
        return rs.MatrixSolverConf(n_outer_iterations=n_outer, solver_cvg_value=cvg,
                                   stop_rule=rs.STOP_RELRES, precision=precision)

    s64 = rs.RMatrixSolver(conf(rs.PRECISION_FP64, args.tol))
    smp = rs.RMatrixSolver(conf(rs.PRECISION_MIXED, args.tol))

    def run(solver):
        t0 = time.perf_counter()
        asm.assemble_heat_device(dk, dq, dT, db)
        asm_ms = asm.last_ms
        zero_x()
        info = solver.solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
        rs.synchronize()
        wall = time.perf_counter() - t0
        return {"wall_s": wall, "assembly_ms": asm_ms, "solve_ms": info.solve_ms,
                "iterations": info.iterations, "true_relres": info.true_relres,
                "reason": rs.REASON_NAMES.get(info.reason), "refinements": info.inner_iterations}

    if args.learned_only:
        return learned_only(rs, args, asm, xyz, dk, dq, dT, db, dx, zero_x, conf)
    # reference solution for the error check and the CPU baseline: the oracle on the full system
    # (untimed here; its own timings are the cpu_baseline object)
    cpu, x_star, xstar_info = None, None, None
    if not args.no_cpu_baseline:
        log("cpu baseline and reference solution: the oracle on the full system")
        cpu, x_star, xstar_info = oracle_leg(rs, xyz, tets, fixed, t_node, k, q, args.tol, args.cpu_reps)
        log(f"cpu baseline: {cpu['value']} s per system (assembly {cpu['assembly_s']} s, solve {cpu['solve_s']} s)")
    else:
        log("--no-cpu-baseline: solution errors against the GPU's FP64 solve to 1e-13 instead of the oracle's")
        sref = rs.RMatrixSolver(conf(rs.PRECISION_FP64, 1e-13))
        asm.assemble_heat_device(dk, dq, dT, db)
        zero_x()
        iref = sref.solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
        x_star = dx.download()
        xstar_info = {"iterations": iref.iterations, "true_relres": iref.true_relres, "source": "gpu fp64"}

    log("aggregation hierarchy (nested coordinate bins) for the learned preconditioner, built on the GPU")
    bins = (max(1, c // 2),) * 3
    rs.synchronize()
    t0 = time.perf_counter()
    L = rs.LearnedPreconditioner.from_bins(asm.matrix, asm, bins, max_coarse=rs.MAX_COARSE)
    rs.synchronize()
    t_hier = time.perf_counter() - t0
    # the host builder of the same level maps: a check of the GPU maps, and its own time
    t0 = time.perf_counter()
    maps = rs.hierarchy_by_bins(xyz, asm.book(), 1, bins, max_coarse=rs.MAX_COARSE)
    t_host_bins = time.perf_counter() - t0
    if len(maps) != L.n_levels - 1 or not all(np.array_equal(L.level_map(l), m) for l, m in enumerate(maps)):
        raise RuntimeError("GPU hierarchy differs from the host builder's level maps")
    del maps
    params, summary = rs.load_weights()
    trained = params is not None
    L.set_params(params if trained else rs.init_params(0))
    L.set_cycle(args.pre, args.post, rs.COARSE_NETWORK)
    L.set_coarse_cycle(*coarse_degrees(args))
    log(f"levels {list(map(int, L.level_rows))} built in {t_hier:.1f}s; network "
        f"{'trained (' + rs.WEIGHTS_PATH + ')' if trained else 'UNTRAINED (no weights file)'}")
    sl = rs.RMatrixSolver(conf(rs.PRECISION_MIXED, args.tol))

    def run_learned():
        t0 = time.perf_counter()
        asm.assemble_heat_device(dk, dq, dT, db)
        asm_ms = asm.last_ms
        zero_x()
        info = sl.solve_learned_device(asm.matrix, L, db, dx)
        rs.synchronize()
        wall = time.perf_counter() - t0
        return {"wall_s": wall, "assembly_ms": asm_ms, "solve_ms": info.solve_ms,
                "iterations": info.iterations, "true_relres": info.true_relres,
                "reason": rs.REASON_NAMES.get(info.reason), "refinements": info.inner_iterations,
                "fallback": info.fallback, "network_applications": info.precond_applications}

    for i in range(args.warmup):
        r64, rmp, rl = run(s64), run(smp), run_learned()
        log(f"warmup {i}: fp64 {r64['wall_s']:.3f}s ({r64['iterations']} its) "
            f"mixed {rmp['wall_s']:.3f}s ({rmp['iterations']} its) "
            f"learned {rl['wall_s']:.3f}s ({rl['iterations']} its)")
    rs.synchronize()
    res64, resmp, resl = [], [], []
    err_mp = err_l = None
    t_start = time.perf_counter()
    for i in range(args.steps):
        r64 = run(s64)
        rmp = run(smp)
        x_mp = dx.download() if i == args.steps - 1 else None
        rl = run_learned()
        res64.append(r64)
        resmp.append(rmp)
        resl.append(rl)
        log(f"step {i}: fp64 {r64['wall_s']:.3f}s ({r64['iterations']} its, relres "
            f"{r64['true_relres']:.2e}) mixed {rmp['wall_s']:.3f}s ({rmp['iterations']} its) "
            f"learned {rl['wall_s']:.3f}s ({rl['iterations']} its, relres {rl['true_relres']:.2e}, "
            f"fallback {rl['fallback']})")
    rs.synchronize()
    t_total = time.perf_counter() - t_start
    x_l = dx.download()
    if x_mp is not None:
        err_mp = float(np.linalg.norm(x_mp - x_star) / np.linalg.norm(x_star))
    err_l = float(np.linalg.norm(x_l - x_star) / np.linalg.norm(x_star))

    def summarize(rr):
        med = lambda key: float(np.median([r[key] for r in rr]))  # noqa: E731
        out = {"time_s": round(med("wall_s"), 4), "assembly_ms": round(med("assembly_ms"), 3),
               "solve_ms": round(med("solve_ms"), 2), "iterations": int(med("iterations")),
               "true_relres": max(r["true_relres"] for r in rr),
               "reason": rr[-1]["reason"], "fp64_residual_replacements": int(med("refinements"))}
        if "fallback" in rr[0]:
            out["fallback"] = max(r["fallback"] for r in rr)
        return out

    m64, mmp, ml = summarize(res64), summarize(resmp), summarize(resl)
    x64 = None
    zero_x()
    asm.assemble_heat_device(dk, dq, dT, db)
    s64.solve_device(asm.matrix, db, dx, rs.PRECOND_JACOBI, 1)
    x64 = dx.download()
    m64["solution_error_vs_oracle_fp64"] = float(np.linalg.norm(x64 - x_star) / np.linalg.norm(x_star))
    mmp["solution_error_vs_oracle_fp64"] = err_mp
    ml["solution_error_vs_oracle_fp64"] = err_l
    ml["network"] = "trained" if trained else "untrained"
    ml["levels"] = [int(v) for v in L.level_rows]
    ml["cycle"] = {"pre": args.pre, "post": args.post, "coarse_pre": coarse_degrees(args)[0],
                   "coarse_post": coarse_degrees(args)[1], "smoother": "Chebyshev in D^-1 A, FP32"}
    ml["hierarchy_build_s"] = round(t_hier, 4)
    ml["hierarchy_build_note"] = ("GPU: nested coordinate bins, members, next-level patterns and gather lists "
                                  "(rs_learned_create_by_bins); host binning of the same maps alone: "
                                  f"{t_host_bins:.3f} s")
    # the hierarchy depends on the mesh only (coordinate bins, gather lists): built once per mesh
    # and reused by every system on it; for a one-off system it adds to the solve
    ml["time_with_hierarchy_build_s"] = round(ml["time_s"] + t_hier, 4)
    best = ml["time_s"]

    # ablation of the learned part: the same solve with the coarsest level by coarse Jacobi
    # (what the untrained network computes) and by the exact dense inverse
    ablation = {"network": {"iterations": ml["iterations"], "time_s": ml["time_s"]}}
    for name, mode in (("coarse_jacobi", rs.COARSE_JACOBI), ("exact_inverse", rs.COARSE_EXACT)):
        L.set_cycle(args.pre, args.post, mode)
        L.set_coarse_cycle(*coarse_degrees(args))
        r = run_learned()
        ablation[name] = {"iterations": r["iterations"], "time_s": round(r["wall_s"], 4),
                          "true_relres": r["true_relres"]}
        log(f"ablation {name}: {r['wall_s']:.3f}s ({r['iterations']} its)")
    L.set_cycle(args.pre, args.post, rs.COARSE_NETWORK)
    L.set_coarse_cycle(*coarse_degrees(args))
    ablation["note"] = ("exact_inverse: the coarsest operator inverted on the GPU once per system (blocked "
                        "sweep operator in FP64, 16 launches) and applied as a dense FP32 product; its "
                        "time includes that inversion")

    # roofline of the dominant kernel (CSR SpMV), stream events on the library stream
    spmv = {}
    for name, prec, sv, sx in (("fp64", rs.PRECISION_FP64, 8, 8), ("fp32", rs.PRECISION_MIXED, 4, 4)):
        plan, tune_ms = asm.matrix.spmv_plan(prec)
        ms = asm.matrix.spmv_timed(prec, reps=200)
        byts = spmv_bytes(n, nnz, sv, sx, rs.SPMV_COLUMN_BYTES[plan])
        spmv[name] = {"avg_us": round(ms * 1e3, 3), "bytes": byts,
                      "gbs": round(byts / (ms * 1e-3) / 1e9, 1),
                      "shape": rs.SPMV_SHAPE_NAMES[plan],
                      "tuning_us": {rs.SPMV_SHAPE_NAMES[i]: round(t * 1e3, 2) for i, t in enumerate(tune_ms)}}
    # the FP64 CG iteration inside the solver (dir, fused SpMV + p.q, update, two finalizes)
    # against its fused minimum: SpMV bytes + 11 x 8 x n (x, p, r, q, D^-1 read; x, r written in the
    # update; D^-1, r, p read and p written in the direction kernel)
    plan64 = asm.matrix.spmv_plan(rs.PRECISION_FP64)[0]
    it_us = m64["solve_ms"] * 1e3 / max(1, m64["iterations"])
    fused = spmv_bytes(n, nnz, 8, 8, rs.SPMV_COLUMN_BYTES[plan64]) + 11 * 8 * n
    cg_iteration = {"us_per_iteration": round(it_us, 2), "fused_minimum_bytes": fused,
                    "gbs": round(fused / (it_us * 1e-6) / 1e9, 1),
                    "frac": round(fused / (it_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                    "note": "solve_ms / iterations of the FP64 Jacobi-CG, set-up included"}
    # the solver's fused FP64 product (q = A p with the partial p.q sums) launched on its own: the
    # same kernel that runs inside the CG iterations, outside their cache context
    ms = asm.matrix.spmv_timed(rs.TIMED_CG_PRODUCT, reps=200)
    spmv["fp64_cg_product"] = {"avg_us": round(ms * 1e3, 3), "bytes": spmv["fp64"]["bytes"],
                               "gbs": round(spmv["fp64"]["bytes"] / (ms * 1e-3) / 1e9, 1),
                               "shape": spmv["fp64"]["shape"]}
    ach = spmv["fp64"]["gbs"]
    traffic = MEASURED_TRAFFIC_128.get(spmv["fp64"]["shape"]) if (c == 128 and args.sigma == 2.0) else None
    roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": int(traffic) if traffic else None,
                "kernel": f"rs::k_spmv FP64 CSR SpMV ({spmv['fp64']['shape']})",
                "algorithmic_bytes": spmv["fp64"]["bytes"]}

    # post-processing (RSolverHeat::process): heat flux per tet from the solved temperatures
    dTn = rs.DeviceArray.from_host(t_node)
    asm.dofs_to_nodes(dx, dTn)
    dflux = rs.DeviceArray(3 * tets.shape[0])
    fl_ms = []
    for _ in range(6):
        asm.heat_flux_device(dk, dTn, dflux)
        fl_ms.append(asm.last_ms)
    fl = float(np.median(fl_ms[1:]))
    fl_bytes = tets.shape[0] * (16 + 4 + 8 + 24) + xyz.shape[0] * (24 + 8)
    post = {"heat_flux_ms": round(fl, 4), "algorithmic_bytes": fl_bytes,
            "gbs": round(fl_bytes / (fl * 1e-3) / 1e9, 1)}
    log(f"heat flux: {fl:.3f} ms ({post['gbs']} GB/s algorithmic)")

    out = {
        "metric": METRIC, "value": round(best, 4), "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_total / max(1, args.steps) * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f64 (plain) / f32 SpMV+precond, f64 residuals (mixed) / f32 cycle, bf16 network on "
                  "MFMA, f64 residuals (learned)"),
        "data": "synthetic",
        "config": {"workload": f"steady heat {c}^3 cells x 6 tets, lognormal(0,{args.sigma}) "
                               f"conductivity (seed {args.seed_k}), to relres {args.tol}: learned+mixed "
                               f"flexible CG vs plain FP64 Jacobi-CG",
                   "nodes": int(xyz.shape[0]), "tets": int(tets.shape[0]), "unknowns": n,
                   "nnz": int(nnz), "colours": asm.n_colours, "parallelism": "single GPU"},
        "methods": {"fp64_jacobi_cg": m64, "mixed_jacobi_cg": mmp, "learned_mixed_fcg": ml,
                    "speedup_mixed_vs_fp64": round(m64["time_s"] / mmp["time_s"], 3),
                    "speedup_learned_vs_fp64": round(m64["time_s"] / ml["time_s"], 3)},
        "learned_ablation": ablation,
        "reference_solution": dict(xstar_info, relres_target=1e-13) if xstar_info else None,
        "spmv": spmv, "roofline": roofline, "cg_iteration_fp64": cg_iteration, "cpu_baseline": cpu,
        "post_processing": post,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
