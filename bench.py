#!/usr/bin/env python3
"""Benchmark of the monolis Krylov path on one MI355X.

Headline (value): FP64 BCSR3 SpMV throughput in GB/s of algorithmic bytes on
the 27-point, 3-DOF-per-node system on 100^3 nodes (3.0 M DOF, 26,463,592
3x3 blocks), data resident in HBM.  A step is one SpMV; K steps are timed with
HIP events on the library stream after W warm-up steps.

Also reported on the same system (extra keys): CG + block-Jacobi to 1e-8 in
FP64 versus the FP32-inner / FP64-refined CG, with iteration counts, the error
of the refined solution against the FP64 one, and the CPU baseline (the
oracle's OpenMP FP64 SpMV and CG on the host cores, a bounded sample).

Further workloads (extra keys, skipped with --no-extras):
  precision        the same system scaled D A D (D log-uniform over [1, 1e3]),
                   FP64 CG vs FP32-refined CG to 1e-10; 128^3 convection-
                   diffusion (cell Peclet 0.5), FP64 vs FP32-refined BiCGSTAB;
  batching         128^3 Poisson with 64 seeded right-hand sides as one
                   multi-RHS CG (FP64 and FP32-refined) vs 64 single solves;
                   512 Poisson systems with edges 10..40 in one persistent,
                   graph-captured launch (autotuned);
  differentiation  64^3 BCSR3 SPD and 64^3 convection-diffusion: forward solve
                   and adjoint (A^T lambda = dL/dx plus the -lambda x^T kernel);
  complex          the headline pattern with complex values K + i 0.5 D (D the
                   diagonal blocks): complex SpMV GB/s and COCG + complex
                   block-Jacobi to 1e-8 (monolis method 13);
  assembly         device assembly of a 64^3-element hexahedral mesh with
                   3 DOF per node: pattern build, 262,144 element matrices
                   (24 x 24) added in element order, Dirichlet conditions on
                   one face; inputs resident in HBM.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BCSR SpMV GB/s (% HBM peak) + CG time-to-1e-8: FP32-refined vs FP64"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "spmv_bcsr3_stencil27_100cube_hbm_traffic.md")


def measured_traffic(grid: int):
    """HBM bytes per FP64 SpMV launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (calibrated with the copy kernel) of the 100^3 workload."""
    if grid != 100 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":")[1]))
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=100, help="nodes per axis (100 -> 3.0 M DOF)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-solve", action="store_true", help="skip the CG comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU SpMV sample")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--variant", type=int, default=-1,
                    help="force this SpMV kernel variant instead of running the autotuner")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the precision / batching / differentiation workloads")
    ap.add_argument("--multi-spmm", action="store_true",
                    help="only time the 64-RHS SpMM of the 128^3 Poisson workload (profiling runs)")
    ap.add_argument("--fp32", action="store_true", help="with --multi-spmm: the FP32 SpMM")
    ap.add_argument("--spmm-config", type=int, default=-1,
                    help="with --multi-spmm: force this SpMM launch configuration (default: the untuned one)")
    ap.add_argument("--assembly", action="store_true",
                    help="only run the device-assembly workload (profiling runs)")
    ap.add_argument("--complex-spmv", action="store_true",
                    help="only time the complex SpMV of the headline pattern (profiling runs)")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the streaming-copy probe (1 GiB read + 1 GiB write per launch)")
    ap.add_argument("--convdiff-only", action="store_true",
                    help="only the 128^3 convection-diffusion FP64 / refined BiCGSTAB solves (profiling runs)")
    ap.add_argument("--inner-tol-sweep", action="store_true",
                    help="refined solves of the convection-diffusion and headline systems over inner tolerances")
    ap.add_argument("--batch-teams", action="store_true",
                    help="batched CG of k identical 40^3 Poisson systems, k = 1 .. 256: one workgroup per "
                         "system against team configurations (JSON line), then exit")
    ap.add_argument("--cg-variant-sweep", action="store_true",
                    help="time the headline FP64 CG once per SpMV kernel variant (JSON line), then exit")
    ap.add_argument("--trace-summary", metavar="CSV",
                    help="summarise a rocprofv3 kernel_trace.csv per kernel and grid size (no GPU), then exit")
    args = ap.parse_args()

    if args.trace_summary:
        print(trace_summary(args.trace_summary))
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (the single-GPU Krylov path); "
              f"--gpus {args.gpus} / WORLD_SIZE {world} is not supported.", file=sys.stderr)
        return 2

    import numpy as np

    from monolis_amd import _lib, api, metrics, workloads as W

    L = _lib.load()
    _lib.check(L.monolis_hip_init(), "monolis_hip_init")

    def copy_probe(reps):
        import ctypes
        n = 1 << 27                                  # 1 GiB of doubles each way
        src, dst = api.DeviceBuffer(n), api.DeviceBuffer(n)
        src.zero()
        ms_c = ctypes.c_double()
        _lib.check(L.monolis_hip_stream_copy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), n, reps,
                                             ctypes.byref(ms_c)), "stream_copy")
        return {"bytes_per_launch": 16 * n, "ms": round(ms_c.value, 5),
                "gbs": round(16 * n / (ms_c.value * 1e-3) / 1e9, 1)}

    def read_probe(reps):
        import ctypes
        n = 1 << 28                                  # 2 GiB of doubles read per launch
        src = api.DeviceBuffer(n)
        src.zero()
        ms_c = ctypes.c_double()
        _lib.check(L.monolis_hip_stream_read(ctypes.c_void_p(src.ptr), n, reps, ctypes.byref(ms_c)),
                   "stream_read")
        return {"bytes_per_launch": 8 * n, "ms": round(ms_c.value, 5),
                "gbs": round(8 * n / (ms_c.value * 1e-3) / 1e9, 1)}

    if args.multi_spmm:
        pm = W.poisson7(128)
        k = 64
        dt = np.float32 if args.fp32 else np.float64
        dmm = api.DeviceMatrix(pm)
        if args.spmm_config >= 0:
            dmm.set_multi_config(args.spmm_config, fp32=args.fp32)
        X = np.random.default_rng(args.seed).uniform(-1, 1, size=(pm.n, k)).astype(dt)
        dxm, dym = api.DeviceBuffer.from_numpy(X), api.DeviceBuffer(pm.n * k, dt)
        for _ in range(args.warmup):
            dmm.matvec_multi(dxm, dym, k)
        t = time.perf_counter()
        for _ in range(args.steps):
            dmm.matvec_multi(dxm, dym, k)
        ms_m = (time.perf_counter() - t) * 1e3 / max(args.steps, 1)
        log(f"multi-RHS SpMM ({'FP32' if args.fp32 else 'FP64'}, 64 RHS, 128^3): {ms_m:.3f} ms per call")
        print(json.dumps({"multi_spmm_ms_per_call": round(ms_m, 4), "steps": args.steps}), flush=True)
        return 0

    if args.assembly:
        print(json.dumps({"assembly": bench_assembly(args)}), flush=True)
        return 0

    if args.complex_spmv:
        print(json.dumps({"complex": bench_complex(args, solve=False)}), flush=True)
        return 0

    if args.convdiff_only or args.inner_tol_sweep:
        print(json.dumps(bench_refine_sweep(args)), flush=True)
        return 0

    if args.batch_teams:
        print(json.dumps({"batch_teams": bench_batch_teams(args)}), flush=True)
        return 0

    if args.copy_only:
        cp = copy_probe(args.steps)
        log(f"stream copy: {cp['ms']:.3f} ms, {cp['gbs']:.0f} GB/s")
        print(json.dumps({"stream_copy": cp}), flush=True)
        return 0

    t0 = time.time()
    m = W.stencil27_bcsr3(args.grid, seed=args.seed)
    xt = W.uniform_vector(m.nndof, seed=args.seed + 1)
    b = W.rhs_from_solution(m, xt)
    log(f"built {m.meta['workload']}: N_b={m.n} nnzb={m.nz} in {time.time() - t0:.1f}s")

    if args.cg_variant_sweep:
        dm = api.DeviceMatrix(m)
        db, xs = api.DeviceBuffer.from_numpy(b), api.DeviceBuffer(m.npndof)
        sweep = []
        for v in range(L.monolis_hip_num_variants()):
            dm.set_variant(v)
            dm.solve(db, xs, api.Prm().quiet())                   # warm-up
            ts, it = [], 0
            for _ in range(3):
                prm = api.Prm().quiet()
                t, _ = _timed(lambda: dm.solve(db, xs, prm), reps=1)
                ts.append(t)
                it = int(prm.Iarray[4])
            sweep.append({"variant": v, "cg_ms": round(min(ts) * 1e3, 3), "iters": it,
                          "ms_per_iter": round(min(ts) * 1e3 / max(it, 1), 4)})
            log(f"variant {v}: CG {min(ts) * 1e3:.2f} ms, {it} iterations")
        print(json.dumps({"cg_variant_sweep": sweep}), flush=True)
        return 0

    dm = api.DeviceMatrix(m)
    info = dm.info()
    dx = api.DeviceBuffer.from_numpy(W.uniform_vector(m.nndof, seed=args.seed + 2))
    dy = api.DeviceBuffer(m.nndof)
    log(f"uploaded; GPU layout {info['bytes_fp64'] / 1e9:.3f} GB ({info['slots']} slots)")

    # --- autotuner (outside the timed region): per-matrix SELL kernel variant ---
    if args.variant >= 0:
        dm.set_variant(args.variant, fp32=False)
        dm.set_variant(args.variant, fp32=True)
        tune = {"best64": args.variant, "best32": args.variant, "ms64": [], "ms32": []}
    else:
        tune = dm.autotune(fp32=True, reps=10)
        log(f"autotune: FP64 variant {tune['best64']} ({min(tune['ms64']) * 1e3:.1f} us), "
            f"FP32 variant {tune['best32']} ({min(tune['ms32']) * 1e3:.1f} us)")

    # --- headline: SpMV, HIP events around exactly K launches -------------------
    _lib.check(L.monolis_hip_synchronize(), "sync")
    ms = dm.time_matvec(dx, dy, warmup=args.warmup, reps=args.steps)
    _lib.check(L.monolis_hip_synchronize(), "sync")
    B64 = metrics.spmv_bytes(m.n, m.nz, 3, 8)
    gbs = B64 / (ms * 1e-3) / 1e9
    log(f"SpMV FP64: {ms * 1e3:.1f} us/launch, {gbs:.0f} GB/s ({100 * gbs / metrics.HBM_PEAK_GBS:.1f} % of peak)")
    # per-launch event pairs (outside the timed batch above): median / spread over >= 50 launches
    each = dm.time_matvec_each(dx, dy, warmup=2, reps=max(50, min(args.steps, 200)))
    med = float(np.median(each))
    per_launch = {"reps": int(len(each)), "median_ms": round(med, 5), "min_ms": round(float(each.min()), 5),
                  "max_ms": round(float(each.max()), 5), "median_gbs": round(B64 / (med * 1e-3) / 1e9, 1)}
    log(f"SpMV FP64 per launch: median {med * 1e3:.1f} us over {len(each)} launches "
        f"(min {each.min() * 1e3:.1f}, max {each.max() * 1e3:.1f})")

    dx32 = api.DeviceBuffer.from_numpy(W.uniform_vector(m.nndof, seed=args.seed + 2).astype(np.float32))
    dy32 = api.DeviceBuffer(m.nndof, np.float32)
    ms32 = dm.time_matvec(dx32, dy32, warmup=args.warmup, reps=args.steps)
    B32 = metrics.spmv_bytes(m.n, m.nz, 3, 4)
    gbs32 = B32 / (ms32 * 1e-3) / 1e9
    log(f"SpMV FP32: {ms32 * 1e3:.1f} us/launch, {gbs32:.0f} GB/s")

    cp = copy_probe(20)
    log(f"stream copy probe: {cp['gbs']:.0f} GB/s")
    rp = read_probe(20)
    log(f"stream read probe: {rp['gbs']:.0f} GB/s; the FP64 SpMV reaches {gbs / rp['gbs']:.3f} of it")

    out = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 27-point BCSR3 SPD system, in-repo generator)",
        "config": {"workload": f"stencil27_bcsr3 {args.grid}^3 nodes x 3 DOF "
                               f"({m.nndof} DOF, {m.nz} 3x3 blocks), FP64 BCSR SpMV",
                   "block": 3, "n_block_rows": m.n, "nnzb": m.nz, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": metrics.HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / metrics.HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.grid),
                     "kernel": "k_sell<double,double,3,OpSpmv>", "algorithmic_bytes": B64},
        "spmv_per_launch": per_launch,
        "spmv_fp32": {"ms": round(ms32, 5), "gbs": round(gbs32, 1), "algorithmic_bytes": B32},
        "stream_copy": cp,
        "stream_read": rp,
        "spmv_vs_read_probe": round(gbs / rp["gbs"], 4),
        "autotune": {"variant_fp64": tune["best64"], "variant_fp32": tune["best32"],
                     "us_fp64": [round(v * 1e3, 1) for v in tune["ms64"]],
                     "us_fp32": [round(v * 1e3, 1) for v in tune["ms32"]]},
    }

    # --- CG FP64 vs FP32-refined on the same system -------------------------------
    if not args.no_solve:
        # the solvers run the SpMV variant that is fastest inside a block-Jacobi CG iteration
        # (timed over whole iterations), which need not be the back-to-back winner above
        stune = {"fp64": dm.autotune_solver(fp32=False, iters=20, reps=2),
                 "fp32": dm.autotune_solver(fp32=True, iters=20, reps=2)}
        log(f"solver-context autotune: FP64 variant {stune['fp64']['best']} "
            f"({min(stune['fp64']['ms_per_iter']):.4f} ms/iteration), FP32 variant {stune['fp32']['best']} "
            f"({min(stune['fp32']['ms_per_iter']):.4f} ms/iteration)")
        tune["solver64"], tune["solver32"] = stune["fp64"]["best"], stune["fp32"]["best"]
        out["autotune"]["solver_variant_fp64"] = stune["fp64"]["best"]
        out["autotune"]["solver_variant_fp32"] = stune["fp32"]["best"]
        out["autotune"]["solver_ms_per_iter_fp64"] = [round(v, 4) for v in stune["fp64"]["ms_per_iter"]]
        out["autotune"]["solver_ms_per_iter_fp32"] = [round(v, 4) for v in stune["fp32"]["ms_per_iter"]]
        db = api.DeviceBuffer.from_numpy(b)
        x64 = api.DeviceBuffer(m.npndof)
        prm = api.Prm().quiet()
        dm.solve(db, x64, prm)                       # warm-up (builds the block-Jacobi factors)
        prm = api.Prm().quiet()
        t = time.perf_counter()
        dm.solve(db, x64, prm)
        t64 = time.perf_counter() - t
        it64 = int(prm.Iarray[4])
        log(f"CG FP64: {it64} iterations, {t64 * 1e3:.1f} ms, resid {prm.Rarray[2]:.2e}")
        xr = api.DeviceBuffer(m.npndof)
        prm_r = api.Prm().quiet()
        dm.solve_refined(db, xr, prm_r, inner_tol=args.inner_tol)   # warm-up (FP32 layout)
        prm_r = api.Prm().quiet()
        t = time.perf_counter()
        st = dm.solve_refined(db, xr, prm_r, inner_tol=args.inner_tol)
        tr = time.perf_counter() - t
        x64h, xrh = x64.numpy(), xr.numpy()
        A = m.to_scipy()
        res64 = float(np.linalg.norm(b - A @ x64h) / np.linalg.norm(b))
        resr = float(np.linalg.norm(b - A @ xrh) / np.linalg.norm(b))
        err = float(np.linalg.norm(xrh - x64h) / np.linalg.norm(x64h))
        errt = float(np.linalg.norm(xrh - xt) / np.linalg.norm(xt))
        log(f"CG refined: {st['outer']} outer / {st['inner_total']} inner, {tr * 1e3:.1f} ms, "
            f"speedup {t64 / tr:.2f}x, rel. err vs FP64 {err:.2e}")
        # monolis's time split (Rarray time_spmv / time_dotp / time_prec) of one more FP64 solve with
        # is_measurement on: an event after every kernel charges each interval to its kernel class
        prm_m = api.Prm().quiet()
        prm_m.Iarray[api.MONOLIS_PRM_I_IS_MEASUREMENT] = 1
        dm.solve(db, api.DeviceBuffer(m.npndof), prm_m)
        split = {"solve_ms": round(prm_m.Rarray[3] * 1e3, 3), "spmv_ms": round(prm_m.Rarray[5] * 1e3, 3),
                 "dotp_ms": round(prm_m.Rarray[6] * 1e3, 3), "prec_ms": round(prm_m.Rarray[7] * 1e3, 3),
                 "iters": int(prm_m.Iarray[4])}
        log(f"CG FP64 time split (is_measurement): SpMV {split['spmv_ms']:.2f} ms, vector/dot "
            f"{split['dotp_ms']:.2f} ms, preconditioner step {split['prec_ms']:.2f} ms of {split['solve_ms']:.2f} ms")
        out["cg"] = {
            "tol": 1e-8, "precond": "block-Jacobi 3x3",
            "fp64_ms": round(t64 * 1e3, 3), "fp64_iters": it64, "fp64_true_resid": res64,
            "fp64_ms_per_iter": round(t64 * 1e3 / max(it64, 1), 4),
            "refined_ms": round(tr * 1e3, 3), "refined_outer": st["outer"],
            "refined_inner_total": st["inner_total"], "refined_true_resid": resr,
            "refined_inner_tol": args.inner_tol,
            "speedup_refined_vs_fp64": round(t64 / tr, 3),
            "refined_rel_err_vs_fp64": err, "refined_rel_err_vs_xtrue": errt,
            "time_split_fp64": split,
            "cg_iter_bytes_fp64": metrics.cg_iter_bytes(m.n, m.nz, 3, 8),
            "cg_fp64_gbs": round(metrics.cg_iter_bytes(m.n, m.nz, 3, 8) * it64 / t64 / 1e9, 1),
        }

        # every real Krylov method on the same system and right-hand side
        names = {1: "CG", 2: "GropCG", 3: "PipeCG", 4: "PipeCR", 5: "BiCGSTAB", 6: "PipeBiCGSTAB",
                 7: "BiCGSTAB_noprec", 8: "PipeBiCGSTAB_noprec"}
        methods = {}
        for meth, name in names.items():
            xm = api.DeviceBuffer(m.npndof)
            dm.solve(db, xm, _prm(method=meth, tol=1e-8))     # warm-up (workspace, factors)
            pm = _prm(method=meth, tol=1e-8)
            tm, _ = _timed(lambda: dm.solve(db, xm, pm), reps=1)
            itm = int(pm.Iarray[4])
            xmh = xm.numpy()
            methods[name] = {"method": meth, "status": int(pm.Iarray[5]), "iters": itm, "ms": round(tm * 1e3, 3),
                             "ms_per_iter": round(tm * 1e3 / max(itm, 1), 4),
                             "true_resid": float(np.linalg.norm(b - A @ xmh) / np.linalg.norm(b))}
            log(f"{name}: {itm} iterations, {tm * 1e3:.1f} ms ({tm * 1e3 / max(itm, 1):.3f} ms/iter), "
                f"true resid {methods[name]['true_resid']:.2e}")
            del xm
        out["methods"] = {"workload": "headline system, block-Jacobi (methods 7 and 8: none) to 1e-8, FP64",
                          **methods}

        # monolis's own calling convention: the flat entry point on host arrays.  Each call
        # stages the matrix (PCIe + SELL build + block-Jacobi factors) and the vectors, then
        # solves on the device and copies x back; this is the PCIe-inclusive figure.
        mono = api.Monolis(m, api.Prm().quiet())
        xh = np.zeros(m.npndof)
        flat = []
        for _ in range(2):                                   # the first call warms up
            mono.prm = api.Prm().quiet()
            t = time.perf_counter()
            api.monolis_solve_R(mono, b, xh)
            flat.append((time.perf_counter() - t, float(mono.prm.Rarray[4]), float(mono.prm.Rarray[3])))
        tw, tprep, tsol = flat[-1]
        h2d = m.nz * 9 * 8 + (m.np_ + 1 + m.nz) * 4 + 2 * m.npndof * 8
        out["flat_abi_host"] = {
            "workload": "monolis_solve_R_c_main on host arrays, headline system, block-Jacobi CG to 1e-8",
            "wall_ms": round(tw * 1e3, 3), "prep_ms": round(tprep * 1e3, 3), "solve_total_ms": round(tsol * 1e3, 3),
            "iters": int(mono.prm.Iarray[4]), "status": int(mono.prm.Iarray[5]),
            "h2d_bytes": int(h2d), "h2d_gbs_over_prep": round(h2d / tprep / 1e9, 2) if tprep > 0 else None,
            "rel_diff_vs_device_solve": float(np.linalg.norm(xh - x64h) / np.linalg.norm(x64h)),
        }
        log(f"flat ABI on host arrays: {tw * 1e3:.1f} ms wall, staging {tprep * 1e3:.1f} ms, "
            f"{int(mono.prm.Iarray[4])} iterations")
        del mono, xh

    # --- CPU baseline: the oracle (monolis's CPU loops, OpenMP) on host cores -----
    if not args.no_cpu:
        import oracle
        cores = os.cpu_count() or 1
        oracle.set_threads(cores)
        xs = W.uniform_vector(m.nndof, seed=args.seed + 2)
        t1 = oracle.time_matvec(m, xs, 1)
        reps = max(1, min(500, int(args.cpu_seconds / max(t1, 1e-6))))
        tc = oracle.time_matvec(m, xs, reps) / reps
        cpu_gbs = B64 / tc / 1e9
        log(f"CPU SpMV ({cores} threads): {tc * 1e3:.1f} ms, {cpu_gbs:.1f} GB/s")
        out["cpu_baseline"] = {"value": round(cpu_gbs, 2), "unit": "GB/s", "cores": oracle.get_threads(),
                               "kind": "port", "cpu_model": cpu_model(),
                               "sample": f"{reps} FP64 SpMVs of the same 100^3 BCSR3 matrix "
                                         f"({tc * 1e3:.1f} ms each)"}
        if not args.no_solve:
            t = time.perf_counter()
            xo, sto, I, R = oracle.solve(m, b, tol=1e-8)
            tcg = time.perf_counter() - t
            out["cpu_baseline"]["cg_fp64_s"] = round(tcg, 3)
            out["cpu_baseline"]["cg_fp64_iters"] = int(I[4])
            out["cg"]["fp64_iters_oracle"] = int(I[4])
            log(f"CPU CG: {int(I[4])} iterations in {tcg:.2f}s")

    if not args.no_extras and not args.no_solve:
        dm = dx = dy = dx32 = dy32 = None             # release the headline buffers first
        out["precision"] = bench_precision(m, xt, args, tune)
        m = None
        out["batching"] = bench_batching(args)
        out["differentiation"] = bench_differentiation(args)
        out["complex"] = bench_complex(args, cpu=not args.no_cpu)
        out["assembly"] = bench_assembly(args, cpu=not args.no_cpu)

    print(json.dumps(out), flush=True)
    return 0


def trace_summary(path: str) -> str:
    """Markdown table of a rocprofv3 --kernel-trace CSV: per (kernel, grid size)
    the dispatch count and the mean / median / min / max duration in us."""
    import csv
    import statistics
    groups = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            name = r["Kernel_Name"].replace("(anonymous namespace)::", "").split("(")[0].strip()
            if name.startswith("void "):
                name = name[5:]
            key = (name[:120], int(r["Grid_Size_X"]) // max(int(r["Workgroup_Size_X"]), 1))
            groups.setdefault(key, []).append(d)
    lines = ["| kernel | workgroups | calls | mean us | median us | min us | max us |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for (name, wg), ds in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"| `{name}` | {wg} | {len(ds)} | {statistics.mean(ds):.1f} | {statistics.median(ds):.1f} | "
                     f"{min(ds):.1f} | {max(ds):.1f} |")
    return "\n".join(lines)


def cpu_model() -> str:
    """Host CPU model name (/proc/cpuinfo), for the CPU baseline."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _prm(method=1, precond=1, tol=1e-8, max_iter=5000):
    from monolis_amd import api
    prm = api.Prm().quiet()
    prm.Iarray[1], prm.Iarray[2], prm.Iarray[3] = method, precond, max_iter
    prm.Rarray[1] = tol
    return prm


def _timed(fn, reps=1):
    """Wall time of the last of `reps` calls (the first ones warm up)."""
    t = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        t = time.perf_counter() - t0
    return t, r


def bench_precision(m, xt, args, tune):
    """FP32-inner / FP64-refined solves on an ill-conditioned twin of the headline
    system (the headline's tuned SpMV variants: same pattern) and on the
    nonsymmetric convection-diffusion problem (tuned for its own pattern)."""
    import numpy as np
    from monolis_amd import api, workloads as W
    res = {}
    t0 = time.time()
    ms_ = W.scale_symmetric(m, 1e3, seed=args.seed + 5)
    b = W.rhs_from_solution(ms_, xt)
    log(f"built D A D twin in {time.time() - t0:.1f}s")
    dm = api.DeviceMatrix(ms_)
    dm.set_variant(tune.get("solver64", tune["best64"]))
    dm.set_variant(tune.get("solver32", tune["best32"]), fp32=True)
    db = api.DeviceBuffer.from_numpy(b)
    x64, xr = api.DeviceBuffer(ms_.npndof), api.DeviceBuffer(ms_.npndof)
    p64 = _prm(tol=1e-10)
    t64, _ = _timed(lambda: dm.solve(db, x64, p64), reps=2)
    it64 = int(p64.Iarray[4])
    pr = _prm(tol=1e-10)
    tr, st = _timed(lambda: dm.solve_refined(db, xr, pr, inner_tol=args.inner_tol), reps=2)
    A = ms_.to_scipy()
    a64, ar = x64.numpy(), xr.numpy()
    res["ill_conditioned_bcsr3"] = {
        "workload": f"D A D of the headline system, D log-uniform [1, 1e3], block-Jacobi CG to 1e-10",
        "fp64_ms": round(t64 * 1e3, 3), "fp64_iters": it64,
        "refined_ms": round(tr * 1e3, 3), "refined_outer": st["outer"], "refined_inner_total": st["inner_total"],
        "speedup": round(t64 / tr, 3),
        "fp64_true_resid": float(np.linalg.norm(b - A @ a64) / np.linalg.norm(b)),
        "refined_true_resid": float(np.linalg.norm(b - A @ ar) / np.linalg.norm(b)),
        "refined_rel_err_vs_fp64": float(np.linalg.norm(ar - a64) / np.linalg.norm(a64)),
        "refined_rel_err_vs_xtrue": float(np.linalg.norm(ar - xt) / np.linalg.norm(xt)),
    }
    log(f"precision/D A D: FP64 {t64 * 1e3:.1f} ms ({it64} it), refined {tr * 1e3:.1f} ms "
        f"({st['outer']} outer / {st['inner_total']} inner), err vs FP64 "
        f"{res['ill_conditioned_bcsr3']['refined_rel_err_vs_fp64']:.2e}")
    del dm, db, x64, xr, ms_, A
    cd = W.convdiff7(128, peclet=0.5)
    b = W.uniform_vector(cd.nndof, seed=args.seed + 6)
    dm = api.DeviceMatrix(cd)
    if args.variant >= 0:
        dm.set_variant(args.variant)
        dm.set_variant(args.variant, fp32=True)
    else:
        ct = dm.autotune(fp32=True, reps=10)
        log(f"precision/convdiff autotune: FP64 variant {ct['best64']}, FP32 variant {ct['best32']}")
    db = api.DeviceBuffer.from_numpy(b)
    x64, xr = api.DeviceBuffer(cd.npndof), api.DeviceBuffer(cd.npndof)
    p64 = _prm(method=5, tol=1e-8)
    t64, _ = _timed(lambda: dm.solve(db, x64, p64), reps=2)
    pr = _prm(method=5, tol=1e-8)
    tr, st = _timed(lambda: dm.solve_refined(db, xr, pr, inner_tol=args.inner_tol), reps=2)
    A = cd.to_scipy()
    a64, ar = x64.numpy(), xr.numpy()
    res["convdiff_128_bicgstab"] = {
        "workload": "7-point convection-diffusion 128^3 (cell Peclet 0.5), diagonal-scaled BiCGSTAB to 1e-8",
        "fp64_ms": round(t64 * 1e3, 3), "fp64_iters": int(p64.Iarray[4]),
        "refined_ms": round(tr * 1e3, 3), "refined_outer": st["outer"], "refined_inner_total": st["inner_total"],
        "speedup": round(t64 / tr, 3),
        "fp64_true_resid": float(np.linalg.norm(b - A @ a64) / np.linalg.norm(b)),
        "refined_true_resid": float(np.linalg.norm(b - A @ ar) / np.linalg.norm(b)),
        "refined_rel_err_vs_fp64": float(np.linalg.norm(ar - a64) / np.linalg.norm(a64)),
    }
    log(f"precision/convdiff: FP64 {t64 * 1e3:.1f} ms, refined {tr * 1e3:.1f} ms")
    return res


def bench_refine_sweep(args):
    """FP64 vs FP32-refined solves at several inner tolerances: BiCGSTAB on the 128^3
    convection-diffusion system and (with --inner-tol-sweep) CG on the headline system.
    With --convdiff-only: one FP64 and one refined BiCGSTAB solve at --inner-tol (after a
    warm-up of each), for kernel traces."""
    from monolis_amd import api, workloads as W
    out = {}
    cd = W.convdiff7(128, peclet=0.5)
    b = W.uniform_vector(cd.nndof, seed=args.seed + 6)
    dm = api.DeviceMatrix(cd)
    dm.set_variant(4)
    dm.set_variant(4, fp32=True)
    db, x = api.DeviceBuffer.from_numpy(b), api.DeviceBuffer(cd.npndof)
    tols = [args.inner_tol] if args.convdiff_only else [1e-4, 1e-5, 1e-6, 1e-7]
    p64 = _prm(method=5, tol=1e-8)
    t64, _ = _timed(lambda: dm.solve(db, x, p64), reps=2)
    rows = []
    for it in tols:
        pr = _prm(method=5, tol=1e-8)
        tr, st = _timed(lambda: dm.solve_refined(db, x, pr, inner_tol=it), reps=2)
        rows.append({"inner_tol": it, "ms": round(tr * 1e3, 3), "outer": st["outer"], "inner": st["inner_total"],
                     "speedup": round(t64 / tr, 3)})
        log(f"convdiff refined, inner tol {it:g}: {tr * 1e3:.2f} ms, {st['outer']} outer / {st['inner_total']} inner")
    out["convdiff_128_bicgstab"] = {"fp64_ms": round(t64 * 1e3, 3), "fp64_iters": int(p64.Iarray[4]), "refined": rows}
    if args.convdiff_only:
        return out
    del dm, db, x
    m = W.stencil27_bcsr3(args.grid, seed=args.seed)
    b = W.rhs_from_solution(m, W.uniform_vector(m.nndof, seed=args.seed + 1))
    dm = api.DeviceMatrix(m)
    dm.autotune(fp32=True, reps=5)
    db, x = api.DeviceBuffer.from_numpy(b), api.DeviceBuffer(m.npndof)
    p64 = _prm(method=1, tol=1e-8)
    t64, _ = _timed(lambda: dm.solve(db, x, p64), reps=2)
    rows = []
    for it in (1e-3, 1e-4, 1e-5, 1e-6):
        pr = _prm(method=1, tol=1e-8)
        tr, st = _timed(lambda: dm.solve_refined(db, x, pr, inner_tol=it), reps=2)
        rows.append({"inner_tol": it, "ms": round(tr * 1e3, 3), "outer": st["outer"], "inner": st["inner_total"],
                     "speedup": round(t64 / tr, 3)})
        log(f"headline refined CG, inner tol {it:g}: {tr * 1e3:.2f} ms, {st['outer']} outer / {st['inner_total']} inner")
    out["headline_cg"] = {"fp64_ms": round(t64 * 1e3, 3), "fp64_iters": int(p64.Iarray[4]), "refined": rows}
    return out


def bench_batching(args):
    """Many right-hand sides (SpMM) and many small systems (one persistent launch)."""
    import numpy as np
    from monolis_amd import api, workloads as W
    res = {}
    k = 64
    pm = W.poisson7(128)
    B = np.random.default_rng(args.seed + 7).uniform(-1, 1, size=(pm.n, k))
    dm = api.DeviceMatrix(pm)
    tune64 = dm.autotune_multi(k, fp32=False, reps=5)
    tune32 = dm.autotune_multi(k, fp32=True, reps=5)
    db = api.DeviceBuffer.from_numpy(B)
    dx = api.DeviceBuffer(pm.n * k)
    p = _prm(tol=1e-8)
    t64, r64 = _timed(lambda: dm.solve_multi(db, dx, k, p), reps=2)
    X64 = dx.numpy().reshape(pm.n, k)
    pr = _prm(tol=1e-8)
    tr, rr = _timed(lambda: dm.solve_multi_refined(db, dx, k, pr, inner_tol=1e-5), reps=2)
    Xr = dx.numpy().reshape(pm.n, k)
    # the same 64 columns as independent single-RHS solves
    b1, x1 = api.DeviceBuffer(pm.n), api.DeviceBuffer(pm.n)
    ps = _prm(tol=1e-8)
    b1.upload(np.ascontiguousarray(B[:, 0]))
    dm.solve(b1, x1, ps)                              # warm-up
    t = time.perf_counter()
    iters1 = []
    for j in range(k):
        b1.upload(np.ascontiguousarray(B[:, j]))
        dm.solve(b1, x1, ps)
        iters1.append(int(ps.Iarray[4]))
    tseq = time.perf_counter() - t
    spmm_bytes = (pm.nz * 12 + (pm.n + 1) * 4) + 2 * pm.n * k * 8
    best_ms = tune64["ms"][tune64["best"]]
    res["multi_rhs_poisson_128"] = {
        "workload": "7-point Poisson 128^3 CSR, 64 seeded uniform RHS, DIAG CG to 1e-8",
        "fp64_ms": round(t64 * 1e3, 2), "fp64_iters_max": int(r64["iters"].max()),
        "fp64_iters_min": int(r64["iters"].min()),
        "refined_ms": round(tr * 1e3, 2), "refined_outer": rr["outer"],
        "refined_inner_max": int(rr["inner_iters"].max()), "refined_inner_tol": 1e-5,
        "refined_rel_err_vs_fp64": float(np.linalg.norm(Xr - X64) / np.linalg.norm(X64)),
        "sequential_64_single_solves_ms": round(tseq * 1e3, 2),
        "single_iters_max": max(iters1),
        "speedup_fp64_vs_sequential": round(tseq / t64, 3),
        "speedup_refined_vs_sequential": round(tseq / tr, 3),
        "spmm_ms": round(best_ms, 4), "spmm_algorithmic_gbs": round(spmm_bytes / (best_ms * 1e-3) / 1e9, 1),
        "spmm_config_fp64": tune64["configs"][tune64["best"]], "spmm_config_fp32": tune32["configs"][tune32["best"]],
    }
    # FP64 iteration: p = z + beta p (3 N k + N), SpMM q = A p with the (p, q) re-read
    # (nnz*12 + (N+1)*4 + 3 N k), x / r / z pass (6 N k + N), 8-byte values
    it_bytes = 12 * pm.n * k * 8 + 2 * pm.n * 8 + pm.nz * 12 + (pm.n + 1) * 4
    it_max = int(r64["iters"].max())
    res["multi_rhs_poisson_128"]["roofline"] = {
        "bound": "hbm", "achieved": round(it_bytes * it_max / t64 / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
        "frac": round(it_bytes * it_max / t64 / 1e9 / 8000.0, 4), "iteration_bytes": it_bytes,
        "ms_per_iteration": round(t64 * 1e3 / it_max, 4),
        "bytes_model": "per FP64 iteration (launches run to the slowest column): 12*N*k*8 + 2*N*8 + nnz*12 + (N+1)*4"}
    log(f"batching/multi-RHS: FP64 {t64 * 1e3:.0f} ms, refined {tr * 1e3:.0f} ms, 64 single solves "
        f"{tseq * 1e3:.0f} ms, SpMM {best_ms:.3f} ms")
    del dm, db, dx, pm, B
    bm, sys_row = W.poisson_batch(512, 10, 40, seed=args.seed + 8)
    bb = W.uniform_vector(bm.n, seed=args.seed + 9)
    dbat = api.DeviceBatch(bm, sys_row)
    db = api.DeviceBuffer.from_numpy(bb)
    dx = api.DeviceBuffer(bm.n)
    p = _prm(tol=1e-8)
    tune = dbat.autotune(db, dx, p, reps=2)
    ms = dbat.time(db, dx, p, reps=5)
    iters, resid, status = dbat.solve(db, dx, p)
    # a sample of the same systems solved one at a time through the single-system path
    sample = list(range(0, 512, 32))
    tsmp = 0.0
    for s_ in sample:
        r0, r1 = int(sys_row[s_]), int(sys_row[s_ + 1])
        j0, j1 = int(bm.index[r0]), int(bm.index[r1])
        from monolis_amd.matrix import Matrix
        sm = Matrix(r1 - r0, r1 - r0, 1, (bm.index[r0:r1 + 1] - j0).astype(np.int32),
                    (bm.item[j0:j1] - r0).astype(np.int32), bm.A[j0:j1].copy()).ensure_dof_index()
        sdm = api.DeviceMatrix(sm)
        sb, sx = api.DeviceBuffer.from_numpy(bb[r0:r1]), api.DeviceBuffer(r1 - r0)
        ps = _prm(tol=1e-8)
        sdm.solve(sb, sx, ps)
        t = time.perf_counter()
        sdm.solve(sb, sx, ps)
        tsmp += time.perf_counter() - t
    est_seq = tsmp * 512 / len(sample)
    # algorithmic bytes: every system moves, per iteration, 12 B per stored entry and per row 4 B of
    # row length, 24 B for the p update, 16 B of SpMV vectors and 64 B for the fused x/r/z pass
    rows_s = np.diff(np.asarray(sys_row, np.int64))
    nnz_s = np.asarray(bm.index, np.int64)[np.asarray(sys_row[1:], np.int64)] - \
        np.asarray(bm.index, np.int64)[np.asarray(sys_row[:-1], np.int64)]
    it_s = np.asarray(iters, np.int64)
    batch_bytes = int(np.sum(it_s * (nnz_s * 12 + rows_s * (4 + 24 + 16 + 64))))
    res["batch_512_poisson"] = {
        "workload": f"512 independent 7-point Poisson systems, edges U{{10..40}}, {bm.n} DOF total, DIAG CG to 1e-8",
        "roofline": {"bound": "hbm", "achieved": round(batch_bytes / (ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(batch_bytes / (ms * 1e-3) / 1e9 / 8000.0, 4),
                     "bytes": batch_bytes,
                     "bytes_model": "sum over systems of iterations x (12 B per entry + 108 B per row)"},
        "ms": round(ms, 3), "all_converged": bool(np.all(status == 0)),
        "iters_min": int(iters.min()), "iters_max": int(iters.max()),
        "config": tune["configs"][tune["best"]], "autotune_ms": [round(v, 3) for v in tune["ms"]],
        "sequential_est_ms": round(est_seq * 1e3, 1),
        "sequential_sample": f"{len(sample)} systems solved one at a time, scaled to 512",
        "speedup_vs_sequential_est": round(est_seq * 1e3 / ms, 2),
    }
    log(f"batching/512 systems: {ms:.2f} ms per graph replay (sequential est {est_seq * 1e3:.0f} ms)")
    return res


def bench_batch_teams(args):
    """k identical 40^3 Poisson systems (DIAG CG to 1e-8) in one batch: the
    one-workgroup configurations 3 (plain vector streams) and 13 (non-temporal)
    against team configurations 7 (split 4, 4 slots per step) and 12 (split 3):
    microseconds per iteration of the batch."""
    from monolis_amd import api, workloads as W
    rows = []
    for k in (1, 4, 16, 32, 64, 128, 256):
        bm, sys_row = W.poisson_batch(edges=[40] * k)
        dbat = api.DeviceBatch(bm, sys_row)
        db = api.DeviceBuffer.from_numpy(W.uniform_vector(bm.n, seed=args.seed + 9))
        dx = api.DeviceBuffer(bm.n)
        row = {"systems": k}
        for cfg in (3, 13, 7, 12):
            sch = dbat.schedule(cfg)
            dbat.set_config(cfg)
            it, _, st = dbat.solve(db, dx, _prm(tol=1e-8))
            ms = dbat.time(db, dx, _prm(tol=1e-8), reps=3)
            row[f"cfg{cfg}"] = {"members_per_system": sch["slots"] // k, "ms": round(ms, 3),
                                "us_per_iteration": round(ms * 1e3 / int(it.max()), 1),
                                "converged": bool((st == 0).all())}
        log(f"batch teams: {k} x 40^3: " + ", ".join(
            f"cfg {c} T={row[f'cfg{c}']['members_per_system']} {row[f'cfg{c}']['us_per_iteration']} us/it"
            for c in (3, 13, 7, 12)))
        rows.append(row)
        del dbat, db, dx
    return rows


def bench_differentiation(args):
    """Forward solve and adjoint (A^T lambda = dL/dx, dL/dA = -lambda x^T) at 64^3."""
    import numpy as np
    from monolis_amd import api, workloads as W
    res = {}
    for name, mat, method, sym in (("bcsr3_spd_64", W.stencil27_bcsr3(64, seed=args.seed + 10), 1, True),
                                   ("convdiff_64", W.convdiff7(64, peclet=0.5), 5, False)):
        mlen = mat.nndof
        rng = np.random.default_rng(args.seed + 11)
        b = rng.uniform(-1, 1, mlen)
        xtg = rng.uniform(-1, 1, mlen)
        dm = api.DeviceMatrix(mat)
        db, dx = api.DeviceBuffer.from_numpy(b), api.DeviceBuffer(mat.npndof)
        p = _prm(method=method, tol=1e-10)
        tf, _ = _timed(lambda: dm.solve(db, dx, p), reps=2)
        itf = int(p.Iarray[4])
        x = dx.numpy()
        dg = api.DeviceBuffer.from_numpy(2.0 * (x - xtg))
        dl, dA = api.DeviceBuffer(mat.npndof), api.DeviceBuffer(len(mat.A))
        pa = _prm(method=method, tol=1e-10)
        t = time.perf_counter()
        dm.adjoint(dx, dg, dl, dA, pa, symmetric=sym)   # first call builds the transpose (nonsymmetric)
        t_first = time.perf_counter() - t
        ta, _ = _timed(lambda: dm.adjoint(dx, dg, dl, dA, pa, symmetric=sym), reps=2)
        res[name] = {
            "forward_ms": round(tf * 1e3, 3), "forward_iters": itf,
            "adjoint_ms": round(ta * 1e3, 3), "adjoint_iters": int(pa.Iarray[4]),
            "adjoint_first_call_ms": round(t_first * 1e3, 3),
            "method": "CG" if method == 1 else "BiCGSTAB", "transpose": "A itself (SPD)" if sym else "device transpose",
            "grad_values": int(len(mat.A)),
        }
        log(f"differentiation/{name}: forward {tf * 1e3:.1f} ms, adjoint {ta * 1e3:.1f} ms "
            f"(first {t_first * 1e3:.1f} ms)")
    return res


def bench_complex(args, solve=True, cpu=False):
    """Complex SpMV and COCG on the headline pattern with complex values."""
    import numpy as np
    from monolis_amd import _lib, api, metrics, workloads as W
    t0 = time.time()
    m = W.stencil27_bcsr3_complex(args.grid, seed=args.seed)
    rng = np.random.default_rng(args.seed + 20)
    log(f"built {m.meta['workload']} in {time.time() - t0:.1f}s")
    dm = api.DeviceMatrixC(m)
    x = rng.uniform(-1, 1, m.nndof) + 1j * rng.uniform(-1, 1, m.nndof)
    dx, dy = api.DeviceBuffer.from_numpy(x), api.DeviceBuffer(m.nndof, np.complex128)
    ms = dm.time_matvec(dx, dy, warmup=args.warmup, reps=args.steps)
    nbytes = metrics.spmv_bytes(m.n, m.nz, 3, 16)
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"complex SpMV: {ms * 1e3:.1f} us/launch, {gbs:.0f} GB/s ({100 * gbs / metrics.HBM_PEAK_GBS:.1f} % of peak)")
    res = {"workload": m.meta["workload"], "spmv_ms": round(ms, 5), "spmv_gbs": round(gbs, 1),
           "spmv_frac_hbm_peak": round(gbs / metrics.HBM_PEAK_GBS, 4), "algorithmic_bytes": nbytes,
           "kernel": "k_sell<zd,zd,3,OpZSpmv>"}
    if solve:
        xt = rng.uniform(-1, 1, m.nndof) + 1j * rng.uniform(-1, 1, m.nndof)
        dxt = api.DeviceBuffer.from_numpy(xt)
        db = api.DeviceBuffer(m.nndof, np.complex128)
        dm.matvec(dxt, db)
        xs = api.DeviceBuffer(m.npndof, np.complex128)
        p = _prm(method=13, tol=1e-8)
        t, _ = _timed(lambda: dm.solve(db, xs, p), reps=2)   # first call builds the complex block-Jacobi
        it = int(p.Iarray[4])
        err = float(np.linalg.norm(xs.numpy() - xt) / np.linalg.norm(xt))
        ib = metrics.cg_iter_bytes(m.n, m.nz, 3, 16)
        res.update({"cocg_ms": round(t * 1e3, 3), "cocg_iters": it, "cocg_status": int(p.Iarray[5]),
                    "cocg_ms_per_iter": round(t * 1e3 / max(it, 1), 4),
                    "cocg_iter_bytes": ib, "cocg_gbs": round(ib * it / t / 1e9, 1),
                    "cocg_rel_err_vs_xtrue": err, "tol": 1e-8, "precond": "complex block-Jacobi 3x3"})
        log(f"COCG: {it} iterations, {t * 1e3:.1f} ms, rel. err {err:.2e}")
    if cpu:
        import oracle
        oracle.set_threads(os.cpu_count() or 1)
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            oracle.matvec_C(m, x)
        tc = (time.perf_counter() - t) / reps
        res["cpu_spmv_ms"] = round(tc * 1e3, 2)
        res["cpu_spmv_gbs"] = round(nbytes / tc / 1e9, 2)
        res["cpu_threads"] = oracle.get_threads()
        log(f"CPU complex SpMV: {tc * 1e3:.1f} ms")
    return res


def _hex_mesh(nx):
    import numpy as np
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    nid = lambda x, y, z: (z * (nx + 1) + y) * (nx + 1) + x
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    elems = np.stack([nid(ex + dx, ey + dy, ez + dz) for dx, dy, dz in corners], axis=1).astype(np.int32)
    return (nx + 1) ** 3, elems


def bench_assembly(args, cpu=False, nx=64, b=3):
    """Device assembly of a hexahedral mesh (inputs resident in HBM)."""
    import ctypes
    import numpy as np
    from monolis_amd import _lib, api
    L = _lib.load()
    n_node, elems = _hex_mesh(nx)
    n_elem, nzm = len(elems), 8 * b
    rng = np.random.default_rng(args.seed + 30)
    R = rng.uniform(-1, 1, size=(n_elem, nzm, nzm))
    Ke = (np.matmul(R, R.transpose(0, 2, 1)) / nzm + 0.01 * np.eye(nzm)).reshape(-1)
    del R
    de, dk = api.DeviceBuffer.from_numpy(elems), api.DeviceBuffer.from_numpy(Ke)
    h = ctypes.c_void_p()

    def build():
        _lib.check(L.monolis_hip_matrix_create_by_simple_mesh(ctypes.byref(h), n_node, 8, b, n_elem,
                                                              ctypes.c_void_p(de.ptr)), "create_by_simple_mesh")

    build()                                          # warm-up (code objects, allocator)
    L.monolis_hip_matrix_destroy(h)
    t0 = time.perf_counter()
    build()
    t_pattern = time.perf_counter() - t0
    info = (ctypes.c_longlong * 8)()
    L.monolis_hip_matrix_info(h, info)
    nz = int(info[2])

    def add():
        _lib.check(L.monolis_hip_matrix_add_elements_R(h, n_elem, 8, ctypes.c_void_p(de.ptr),
                                                       ctypes.c_void_p(dk.ptr)), "add_elements")

    add()                                            # warm-up
    t0 = time.perf_counter()
    add()
    t_add = time.perf_counter() - t0
    face = np.arange((nx + 1) ** 2, dtype=np.int32)
    nodes = np.repeat(face, b)
    dofs = np.tile(np.arange(b, dtype=np.int32), len(face))
    dn, dd = api.DeviceBuffer.from_numpy(nodes), api.DeviceBuffer.from_numpy(dofs)
    dv = api.DeviceBuffer.from_numpy(np.zeros(len(nodes)))
    rhs = api.DeviceBuffer.from_numpy(np.ones(n_node * b))
    t0 = time.perf_counter()
    _lib.check(L.monolis_hip_set_Dirichlet_bc_R(h, ctypes.c_void_p(rhs.ptr), len(nodes), ctypes.c_void_p(dn.ptr),
                                                ctypes.c_void_p(dd.ptr), ctypes.c_void_p(dv.ptr)), "dirichlet")
    t_bc = time.perf_counter() - t0
    L.monolis_hip_matrix_destroy(h)
    contrib = n_elem * nzm * nzm
    res = {"workload": f"hex8 mesh {nx}^3 elements, {n_node} nodes x {b} DOF, {n_elem} element matrices "
                       f"{nzm}x{nzm}", "nnzb": nz,
           "pattern_ms": round(t_pattern * 1e3, 3), "add_elements_ms": round(t_add * 1e3, 3),
           "add_elements_contributions_per_s": round(contrib / t_add, 1),
           "add_elements_ke_gbs": round(contrib * 8 / t_add / 1e9, 1),
           "dirichlet_ms": round(t_bc * 1e3, 3), "dirichlet_conditions": len(nodes)}
    log(f"assembly: pattern {t_pattern * 1e3:.1f} ms, {n_elem} elements {t_add * 1e3:.1f} ms "
        f"({contrib / t_add / 1e9:.2f} G contributions/s), Dirichlet {t_bc * 1e3:.1f} ms")
    if cpu:
        import oracle
        from monolis_amd import matrix as mm
        t0 = time.perf_counter()
        index, item = oracle.pattern_by_simple_mesh(n_node, elems)
        tp = time.perf_counter() - t0
        mat = mm.Matrix(n_node, n_node, b, index, item, np.zeros(len(item) * b * b))
        sample = n_elem // 8
        t0 = time.perf_counter()
        oracle.add_elements(mat, elems[:sample], Ke[: sample * nzm * nzm])
        ta = (time.perf_counter() - t0) * n_elem / sample
        res["cpu_pattern_ms"] = round(tp * 1e3, 1)
        res["cpu_add_elements_ms_est"] = round(ta * 1e3, 1)
        res["cpu_sample"] = f"pattern of the whole mesh; element add timed on {sample} elements, scaled"
        log(f"CPU assembly: pattern {tp * 1e3:.0f} ms, elements {ta * 1e3:.0f} ms (estimated)")
    return res


if __name__ == "__main__":
    sys.exit(main())
