#!/usr/bin/env python3
"""Benchmark of the sparse block LU factorise-and-solve path on one MI355X.

One step = one factorisation of the workload's matrix from its input tiles
resident in HBM, plus the solve of its right-hand side(s) from b resident in
HBM, to a normwise backward error <= tol in both forms
||b-Ax||_inf/(||A||_inf ||x||_inf + ||b||_inf) and ||b-Ax||_2/(||A||_F ||x||_2 + ||b||_2):

  fp32_ir : FP32 tiles, FP32 MFMA block GEMMs, then FP64 iterative refinement
            (the headline `value`)
  fp64    : the plain FP64 drop-in (FP64 MFMA GEMMs, FP64 tiles)

The default workload is the 1M-unknown 2D convection-diffusion problem (c5) at
tol 1e-10, the configuration the metric is quoted on.

Host work (Matrix Market parse, GPS ordering, planning, tile upload) happens
before the timed region, like the reference's separate "re Order time" /
"plan time" lines.  The reference's own CPU path (built from its sources by
oracle/Makefile) is timed on the same host as the CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sparse-operator-graph-lu_amd"))

import sogl  # noqa: E402
import workloads as W  # noqa: E402

METRIC = "factor+solve s & block-GEMM TFLOP/s at rel. residual ≤1e-10, FP32+IR vs FP64"
PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense MFMA TFLOP/s
HBM_PEAK = 8000.0  # GB/s
# HBM bytes per launch of the FP32 factor kernel in the timed runs, from two rocprofv3
# --pmc runs (FETCH_SIZE and WRITE_SIZE, scaled by what they read for a plain 1 GiB
# copy): profiles/k_factor_persistent_c2_pmc_traffic.md (the timed run: dense tiles)
# and profiles/k_factor_persistent_c5_pmc_traffic.md (16x16 block masks on)
TRAFFIC_FP32_FACTOR = {"c2": 22.46e9, "c5": 24.43e9}
# MFMA busy share of the FP32 factor kernel (SQ_VALU_MFMA_BUSY_CYCLES over the kernel's
# SIMD cycles) and the MFMA flops it executes, from profiles/k_factor_persistent_*_mfma_busy.md:
# at C5 the block masks skip 97% of the nominal products (84% of the 16x16 blocks are zero)
MFMA_FP32_FACTOR = {"c2": {"busy_frac": 0.0586, "executed_gflop": 527.3},
                    "c5": {"busy_frac": 0.0008, "executed_gflop": 46.7}}

WORKLOADS = {
    "c1": ("2D 5-point Poisson 64x64 (n=4,096), 1 RHS U[-1,1] seed 0", 1e-12),
    "c2": ("3D 7-point Laplacian 48^3 (n=110,592), off-diagonals -(1+0.2u) seed 1, diagonal "
           "1.05*sum|off|, nonsymmetric, 1 RHS U[-1,1] seed 1001", 1e-12),
    "c5": ("2D convection-diffusion 1000x1000 (n=1,000,000), 5-point upwind, v=(1,0.5), cell "
           "Peclet 10, 1 RHS U[-1,1] seed 4", 1e-10),
    "c3": ("random nonsymmetric n=16,384 (the largest size whose fill fits the planner; n=262,144 "
           "fills to ~550 GB), band half-width 32 + 4 random entries per row seed 2, N(0,1) values, "
           "diagonal sum|a_ij|+1, 128 RHS N(0,1) seed 1002", 1e-12),
    "sym": ("symmetric input (L L^T path): 3D 7-point variable-coefficient diffusion 48^3 (n=110,592), "
            "log-normal (sigma 1) node coefficients seed 23, lower triangle in the file, 1 RHS U[-1,1] "
            "seed 1023", 1e-12),
}
# --exec choices: (executor mode, resident workgroups per CU; 0 = as many as fit)
EXECUTORS = {"auto": (sogl.EXEC_AUTO, 0), "persistent": (sogl.EXEC_PERSISTENT, 0),
             "persistent-1wg": (sogl.EXEC_PERSISTENT, 1), "levels": (sogl.EXEC_LEVELS, 0)}
# workloads beyond the named configurations, and the fixture holding their reference x
EXTRA = {"sym": (lambda: W.symmetric_lower(W.diffusion_varcoef(48, 3, 23)), "sym_diffusion3d_48")}


def compare_executors(dev, plan, per_op: bool, reuse_mode: int):
    """FP64 factor time of every executor on one resident factor (ms, best of 2), the
    autotuner's choice, and the per-op executor (one launch per unit op) once (the
    factor is loaded without slot reuse, which the per-op executor needs)."""
    dev.set_pool_reuse(sogl.REUSE_OFF)
    fid = dev.load(plan, sogl.FP64)
    dev.set_pool_reuse(reuse_mode)
    out = {}
    dev.factor(fid, sogl.EXEC_GRAPH)  # capture outside the timed runs
    for name, mode in (("persistent", sogl.EXEC_PERSISTENT), ("levels", sogl.EXEC_LEVELS),
                       ("level_graph", sogl.EXEC_GRAPH)):
        ts = []
        for _ in range(2):
            dev.factor(fid, mode)
            ts.append(dev.stats(fid).exec_ms)
        out[name + "_ms"] = round(min(ts), 3)
    best, times = dev.autotune(fid, 2)
    out["autotune_ms"] = {k: round(float(t), 3) for k, t in zip(sogl.TUNE_CANDIDATES, times)}
    out["autotune_choice"] = sogl.TUNE_CANDIDATES[best]
    if per_op:
        t0 = time.perf_counter()
        dev.factor(fid, sogl.EXEC_PER_OP)
        out["per_op_ms"] = round(dev.stats(fid).exec_ms, 3)
        out["per_op_wall_s"] = round(time.perf_counter() - t0, 3)
        out["per_op_launches"] = int(plan.info.n_ops)
    dev.release(fid)
    return out


def cpu_reference(prob, threads: int, timeout: float = 600.0):
    """Time the reference CLI (built from its sources into oracle/_ref) on the workload."""
    exe = os.path.join(ROOT, "oracle", "_ref", "solve")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "w.mtx")
        W.write_mtx(prob, path)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.perf_counter()
        r = subprocess.run([exe, path], capture_output=True, text=True, env=env, timeout=timeout)
        wall = time.perf_counter() - t0
        # the reference's own x at full precision, from the fixture harness linked
        # against the same reference objects (the CLI writes only 6 digits of y)
        xref = None
        dump = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
        if r.returncode == 0 and os.path.exists(dump):
            out = os.path.join(d, "dump")
            os.makedirs(out)
            rd = subprocess.run([dump, path, out], capture_output=True, text=True, env=env, timeout=timeout)
            xf = os.path.join(out, "xorig.f64")
            if rd.returncode == 0 and os.path.exists(xf):
                xref = np.fromfile(xf, dtype=np.float64)[:prob.n]
    if r.returncode != 0:
        return None
    grab = lambda pat: float(re.search(pat, r.stdout).group(1))  # noqa: E731
    return dict(factor_s=grab(r"kernel time: ([0-9.eE+-]+)"), solve_s=grab(r"solve triangled :([0-9.eE+-]+)"),
                reorder_s=grab(r"re Order time: ([0-9.eE+-]+)"), plan_s=grab(r"plan time: ([0-9.eE+-]+)"),
                max_rhs_error=grab(r"max rhs error:([0-9.eE+-]+)"), wall_s=wall, xref=xref)


def physical_cores() -> int:
    """Physical cores this process may run on (its CPU affinity, counted once per core)."""
    try:
        allowed = os.sched_getaffinity(0)
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True).stdout
        cores = set()
        for ln in out.splitlines():
            if ln and not ln.startswith("#"):
                cpu, core, sock = ln.split(",")[:3]
                if int(cpu) in allowed:
                    cores.add((sock, core))
        return max(1, len(cores))
    except (OSError, ValueError):
        return os.cpu_count() or 1


def cpu_description() -> dict:
    """CPU model, sockets, physical cores and logical CPUs of the host, and how many of
    them this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    host_cores, sockets = None, None
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True).stdout
        rows = [ln.split(",") for ln in out.splitlines() if ln and not ln.startswith("#")]
        host_cores = len({(r[2], r[1]) for r in rows})
        sockets = len({r[2] for r in rows})
    except (OSError, IndexError):
        pass
    return {"cpu_model": model, "sockets": sockets, "host_physical_cores": host_cores,
            "host_logical_cpus": os.cpu_count(), "usable_logical_cpus": len(os.sched_getaffinity(0)),
            "usable_physical_cores": physical_cores()}


def run_path(dev, plan, prob, precision, tol, max_iter, steps, warmup, mode, wg_per_cu=0, sweep=(), chain=True):
    """Timed factor + refined-solve steps of one precision.  sweep: numbers of refinement
    corrections after which x is also kept (untimed, after the timed steps).  chain:
    dynamic chaining of the persistent executor."""
    fid = dev.load(plan, precision)
    dev.set_chaining(fid, chain)
    dev.attach_refine(fid, plan)
    n = prob.n
    b = np.asfortranarray(prob.b.reshape(n, -1))
    nrhs = b.shape[1]
    d_b = dev.alloc(b.nbytes)
    d_x = dev.alloc(b.nbytes)
    dev.upload(d_b, b.T.copy())  # column-major n x nrhs
    fac, exe, sol, its, res = [], [], [], [], []
    tuned = None
    if mode == sogl.EXEC_AUTO:  # tune the executor on this factor before warmup, untimed
        best, _ = dev.autotune(fid, 2)
        tuned = sogl.TUNE_CANDIDATES[best]
    for _ in range(warmup):
        dev.factor_config(fid, mode, 1, wg_per_cu)
        dev.solve_refined_device(fid, d_b, d_x, nrhs, tol, max_iter)
    dev.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.factor_config(fid, mode, 1, wg_per_cu)
        st = dev.stats(fid)
        fac.append(st.factor_ms)
        exe.append(st.exec_ms)
        it, r, rc = dev.solve_refined_device(fid, d_b, d_x, nrhs, tol, max_iter)
        sol.append(dev.stats(fid).solve_ms)
        its.append(it)
        res.append(r)
    dev.sync()
    wall = (time.perf_counter() - t0) / steps
    x = dev.download(d_x, np.empty((nrhs, n))).T
    st = dev.stats(fid)
    swept = []
    for k in sweep:  # tol 0: exactly k corrections
        dev.solve_refined_device(fid, d_b, d_x, nrhs, 0.0, k)
        xk = dev.download(d_x, np.empty((nrhs, n))).T
        swept.append((k, xk[:, 0] if nrhs == 1 else xk))
    dev.free(d_b)
    dev.free(d_x)
    dev.release(fid)
    return dict(step_s=wall, factor_s=float(np.mean(fac)) / 1e3, exec_s=float(np.mean(exe)) / 1e3,
                solve_s=float(np.mean(sol)) / 1e3, ir_iters=int(max(its)), residual=float(max(res)),
                pool_gb=st.pool_bytes / 1e9, n_tasks=int(st.n_tasks), n_levels=int(st.n_levels), executor=tuned,
                growth_fp64_tiles=int(st.growth_tiles), zero_block_share=float(st.zero_blocks),
                claimed_tasks=int(st.claimed_tasks), tasks_run=int(st.tasks_run),
                block_masks=bool(st.block_masks), x=x[:, 0] if nrhs == 1 else x, swept=swept)


def calibrate_copy(nbytes: int = 1 << 30, reps: int = 3):
    """A plain elementwise copy of known size: each launch reads and writes nbytes.

    Its FETCH_SIZE / WRITE_SIZE per dispatch give the scale between those counters
    and bytes moved to or from HBM."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": reps}))


def run_serving(args):
    """The serving stream (2D/3D variable-coefficient diffusion, 1-16 RHS, 30% reuse)
    on one GPU: p50/p99 request latency and solves/s."""
    import serve
    reqs = serve.make_stream(args.requests, args.rate)
    m, reqs = serve.serve(reqs, args.slots, sogl.FP64, progress=25)
    worst = 0.0
    for r in reqs:
        prob = serve.matrix_problem(r.matrix)
        worst = max(worst, sogl.oracle_residual(prob.n, prob.row, prob.col, prob.val, r.x, r.b))
    line = {
        "metric": "serving: p50/p99 request latency (ms) and solves/s",
        "value": round(m["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"c4: {args.requests} requests (sample of the 10k stream), Poisson arrivals "
                               f"{args.rate}/s seed 3, 2D side U{{32..512}} / 3D side U{{16..64}} "
                               "variable-coefficient diffusion (log-normal sigma 1), 1-16 RHS, 30% reuse",
                   "slots": args.slots},
        "serving": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in m.items()},
        "max_oracle_residual": worst,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5", choices=sorted(WORKLOADS) + ["c4"],
                    help="c1..c5: the named workloads; sym: a symmetric-input (L L^T) workload")
    ap.add_argument("--requests", type=int, default=200, help="c4: requests in the serving stream")
    ap.add_argument("--rate", type=float, default=20.0, help="c4: mean arrivals per second")
    ap.add_argument("--slots", type=int, default=8, help="c4: concurrent requests on the GPU")
    ap.add_argument("--exec", default="auto", choices=sorted(EXECUTORS),
                    help="auto: the executor configuration the autotuner picks on each factor "
                         "(tuned before the warmup steps, outside the timed region); "
                         "persistent-1wg: the persistent executor with one workgroup per CU")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--pool-reuse", type=int, default=sogl.REUSE_AUTO, choices=(0, 1, 2),
                    help="1: factors reuse pool slots of values no longer read (smaller pool, coherent "
                         "tile loads); 0: one slot per value; 2 (default): reuse when the values would "
                         "take over a quarter of the device's HBM")
    ap.add_argument("--task-links", type=int, default=sogl.LINK_PANEL, choices=(0, 1, 2, 3),
                    help="mask of sogl_ctx_set_task_links: 2 (default) = the lu task runs its most "
                         "critical panel GEMM itself; 1 = after-links (measured slower); 0 = none")
    ap.add_argument("--no-reuse-leg", action="store_true",
                    help="skip the FP64 leg that factors with pool slot reuse forced on")
    ap.add_argument("--no-chain", action="store_true",
                    help="persistent executor without dynamic chaining (for comparison)")
    ap.add_argument("--chain-claims", type=int, default=0,
                    help="dynamic chaining: tasks one finished task may claim for its workgroup")
    ap.add_argument("--bf16", action="store_true",
                    help="also run the BF16 experiment (bf16 MFMA operands, FP32 tiles, FP64 refinement); "
                         "on by default for c5")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16 experiment leg of c5")
    ap.add_argument("--no-ir-sweep", action="store_true",
                    help="skip the refinement sweep (x error of FP32+IR after 0, 1, 2, ... corrections)")
    ap.add_argument("--compare-executors", action="store_true",
                    help="also time every executor (FP64) incl. one launch per unit op (default for c3)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy kernel 3 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the factor on a single MI355X "
              "(multi-GPU sharding is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.config == "c4":
        run_serving(args)
        return
    mode = EXECUTORS[args.exec]
    desc, tol = WORKLOADS[args.config]
    prob = EXTRA[args.config][0]() if args.config in EXTRA else W.config(args.config)
    t0 = time.perf_counter()
    plan = sogl.Plan(prob.n, prob.row, prob.col, prob.val, prob.b if prob.b.ndim == 1 else prob.b[:, 0],
                     symmetric=prob.symmetric)
    t_plan = time.perf_counter() - t0
    info = plan.info
    gemm_flops = (info.n_mul + info.n_mulneg + info.n_mult) * 2.0 * 64 ** 3
    nrhs = 1 if prob.b.ndim == 1 else prob.b.shape[1]
    solve_bytes = {k: (info.n_l_tiles + info.n_u_tiles) * 64 * 64 * s + 2 * plan.N * 8 * nrhs
                   for k, s in (("fp64", 8), ("fp32", 4))}
    dev = sogl.Device(0)
    dev.set_pool_reuse(args.pool_reuse)
    dev.set_task_links(args.task_links)
    paths = {}
    sweep = () if args.no_ir_sweep or nrhs != 1 else tuple(range(0, 7))
    ch = 0 if args.no_chain else args.chain_claims
    paths["fp32_ir"] = run_path(dev, plan, prob, sogl.FP32, tol, 30, args.steps, args.warmup, *mode, sweep=sweep,
                                chain=ch)
    paths["fp64"] = run_path(dev, plan, prob, sogl.FP64, tol, 0, args.steps, args.warmup, *mode, chain=ch)
    if args.bf16 or (args.config == "c5" and not args.no_bf16):
        paths["bf16_ir"] = run_path(dev, plan, prob, sogl.BF16, tol, 50, min(args.steps, 3), min(args.warmup, 1),
                                    *mode, chain=ch)
    reuse_leg = None
    if not args.no_reuse_leg and args.pool_reuse != sogl.REUSE_ON:
        # the same FP64 factor+solve with slot reuse forced on: its pool, its time, and
        # its bits against the leg above
        dev.set_pool_reuse(sogl.REUSE_ON)
        r = run_path(dev, plan, prob, sogl.FP64, tol, 0, min(args.steps, 2), min(args.warmup, 1), *mode, chain=ch)
        dev.set_pool_reuse(args.pool_reuse)
        xr = r.pop("x")
        reuse_leg = {"precision": "fp64", "exec_s": round(r["exec_s"], 6), "step_s": round(r["step_s"], 6),
                     "pool_gb": round(r["pool_gb"], 3), "pool_gb_main_leg": round(paths["fp64"]["pool_gb"], 3),
                     "x_bits_equal_main_leg": bool(np.array_equal(xr, paths["fp64"]["x"]))}
    executors = None
    if args.compare_executors or args.config == "c3":
        executors = compare_executors(dev, plan, per_op=True, reuse_mode=args.pool_reuse)
    lu_same = None
    if prob.symmetric:  # the same matrix given in full (both triangles) takes the LU path
        r, c, v = W.full_matrix(prob)
        full = W.Problem(prob.name + "_full", prob.n, r, c, v, prob.b)
        lu_plan = sogl.Plan(full.n, full.row, full.col, full.val, full.b)
        lu_same = run_path(dev, lu_plan, full, sogl.FP64, tol, 0, args.steps, args.warmup, *mode)
        lu_same.pop("x")
        lu_same = {k: (float(f"{v:.6g}") if isinstance(v, float) else v) for k, v in lu_same.items()}
        lu_same["unit_ops"] = int(lu_plan.info.n_ops)
        del lu_plan
    dev.close()

    ref = None
    if not args.no_cpu:
        threads = min(16, physical_cores())
        ref = cpu_reference(prob if nrhs == 1 else dataclasses.replace(prob, b=prob.b[:, 0].copy()), threads)

    # parity: vs the reference's FP64 x (live reference run, else the golden fixture)
    # and the FP64 oracle residual
    gold = os.path.join(ROOT, "tests", "golden", f"{EXTRA[args.config][1] if args.config in EXTRA else args.config}.npz")
    mat = W.full_matrix(prob)  # the whole matrix (a symmetric input stores one triangle)
    g = np.load(gold, allow_pickle=False) if os.path.exists(gold) else None
    def x_error(x):
        x0 = x if x.ndim == 1 else x[:, 0]
        if ref is not None and ref["xref"] is not None:
            return float(np.linalg.norm(x0 - ref["xref"]) / np.linalg.norm(ref["xref"]))
        if g is not None and "xorig" in g:
            return float(np.linalg.norm(x0 - g["xorig"]) / np.linalg.norm(g["xorig"]))
        if g is not None and "xorig_sample" in g:
            idx = g["xorig_idx"]
            return float(np.linalg.norm(x0[idx] - g["xorig_sample"]) / np.linalg.norm(g["xorig_sample"]))
        return None

    out_paths = {}
    sweeps = {}
    for k, p in paths.items():
        x = p.pop("x")
        resid = sogl.oracle_residual(prob.n, mat[0], mat[1], mat[2], x, prob.b)
        berr = sogl.backward_error_inf(prob.n, mat[0], mat[1], mat[2], x, prob.b)
        err = x_error(x)
        sweeps[k] = [{"corrections": it, "rel_err_vs_reference_x": x_error(xk),
                      "backward_error_inf": sogl.backward_error_inf(prob.n, mat[0], mat[1], mat[2], xk, prob.b),
                      "oracle_residual": sogl.oracle_residual(prob.n, mat[0], mat[1], mat[2], xk, prob.b)}
                     for it, xk in p.pop("swept")]
        prec = "fp64" if k == "fp64" else "fp32"
        p.update(oracle_residual=resid, backward_error_inf=berr, rel_err_vs_reference_x=err,
                 gemm_tflops=gemm_flops / p["exec_s"] / 1e12,
                 gemm_frac_of_peak=gemm_flops / p["exec_s"] / 1e12 / PEAK[prec],
                 solve_gb_s=solve_bytes[prec] * (1 + p["ir_iters"]) / p["solve_s"] / 1e9)
        p["solve_frac_hbm"] = p["solve_gb_s"] / HBM_PEAK
        out_paths[k] = {kk: (float(f"{v:.6g}") if isinstance(v, float) else v) for kk, v in p.items()}

    head = out_paths["fp32_ir"]
    cpu = None
    if ref is not None:
        threads = min(16, physical_cores())
        cpu = {"value": round(ref["factor_s"] + ref["solve_s"], 4), "unit": "s", "cores": threads,
               "kind": "reference",
               "sample": f"the full {args.config} workload, one factor+solve of the reference CLI "
                         f"(built from its sources) at OMP_NUM_THREADS={threads}; value = its "
                         "'kernel time' + 'solve triangled'" +
                         ("" if nrhs == 1 else f" for the first of the {nrhs} RHS (the CLI solves one)"),
               "factor_s": ref["factor_s"], "solve_s": ref["solve_s"],
               "reorder_s": ref["reorder_s"], "plan_s": ref["plan_s"],
               "max_rhs_error": ref["max_rhs_error"], "wall_s": round(ref["wall_s"], 3),
               "all_cores": "the reference runs at most 16 OpenMP threads: its per-thread buffers are "
                            "arrays of MAXTHREAD = 16 (const.h:23, memutil.cpp:33-47) indexed by the "
                            "thread id, so more threads write past them; cores = min(16, usable "
                            "physical cores)"}
        cpu.update(cpu_description())
    sweep = sweeps.get("fp32_ir") or []
    fp64_err = out_paths["fp64"]["rel_err_vs_reference_x"]
    reach = None
    if sweep and fp64_err is not None:
        for e in sweep:
            if e["rel_err_vs_reference_x"] is not None and e["rel_err_vs_reference_x"] <= 10 * fp64_err:
                reach = e["corrections"]
                break
    line = {
        "metric": METRIC,
        # the whole step: factor, then the refined solve (triangular solves, FP64 residuals
        # and corrections, permutations) to the tolerance, inputs resident in HBM
        "value": round(head["step_s"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["step_s"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor + f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {desc}", "tol": tol,
                   "executor": args.exec if head["executor"] is None else f"autotuned: {head['executor']}",
                   "n": prob.n, "nnz": prob.nnz, "unit_ops": int(info.n_ops),
                   "gemm_ops": int(info.n_mul + info.n_mulneg + info.n_mult), "stages": int(info.n_stages),
                   "gemm_gflop": round(gemm_flops / 1e9, 3), "host_plan_s": round(t_plan, 3),
                   "task_links": args.task_links},
        "roofline": {"bound": "mfma", "achieved": round(head["gemm_tflops"], 4), "peak": PEAK["fp32"],
                     "unit": "TFLOP/s", "frac": round(head["gemm_tflops"] / PEAK["fp32"], 6),
                     "traffic": TRAFFIC_FP32_FACTOR.get(args.config),
                     "mfma_pmc": MFMA_FP32_FACTOR.get(args.config),
                     "kernel": "k_factor_persistent<float> (FP32 factor executor); achieved = "
                               "(mul+mulneg+mult ops) x 2*64^3 flop / executor time from HIP events"},
        "cpu_baseline": cpu,
        "fp32_ir": out_paths["fp32_ir"],
        "fp64": out_paths["fp64"],
    }
    if sweep:
        line["fp32_ir_sweep"] = {
            "what": "FP32 factor + k FP64 corrections (untimed, after the timed steps): x error against "
                    "the reference's FP64 x, and both backward-error forms",
            "steps": [{kk: (float(f"{v:.4g}") if isinstance(v, float) else v) for kk, v in e.items()} for e in sweep],
            "corrections_to_reach_10x_fp64_x_error": reach}
    if executors is not None:
        line["executors_fp64"] = executors
    if reuse_leg is not None:
        line["pool_reuse_fp64"] = reuse_leg
    if "bf16_ir" in out_paths:
        line["bf16_ir_experiment"] = out_paths["bf16_ir"]
    if lu_same is not None:
        line["fp64_lu_same_matrix"] = lu_same
    print(json.dumps(line))


if __name__ == "__main__":
    main()
