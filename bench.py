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
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the streaming-copy probe (1 GiB read + 1 GiB write per launch)")
    args = ap.parse_args()

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

    dx32 = api.DeviceBuffer.from_numpy(W.uniform_vector(m.nndof, seed=args.seed + 2).astype(np.float32))
    dy32 = api.DeviceBuffer(m.nndof, np.float32)
    ms32 = dm.time_matvec(dx32, dy32, warmup=args.warmup, reps=args.steps)
    B32 = metrics.spmv_bytes(m.n, m.nz, 3, 4)
    gbs32 = B32 / (ms32 * 1e-3) / 1e9
    log(f"SpMV FP32: {ms32 * 1e3:.1f} us/launch, {gbs32:.0f} GB/s")

    cp = copy_probe(20)
    log(f"stream copy probe: {cp['gbs']:.0f} GB/s")

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
        "spmv_fp32": {"ms": round(ms32, 5), "gbs": round(gbs32, 1), "algorithmic_bytes": B32},
        "stream_copy": cp,
        "autotune": {"variant_fp64": tune["best64"], "variant_fp32": tune["best32"],
                     "us_fp64": [round(v * 1e3, 1) for v in tune["ms64"]],
                     "us_fp32": [round(v * 1e3, 1) for v in tune["ms32"]]},
    }

    # --- CG FP64 vs FP32-refined on the same system -------------------------------
    if not args.no_solve:
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
        out["cg"] = {
            "tol": 1e-8, "precond": "block-Jacobi 3x3",
            "fp64_ms": round(t64 * 1e3, 3), "fp64_iters": it64, "fp64_true_resid": res64,
            "fp64_ms_per_iter": round(t64 * 1e3 / max(it64, 1), 4),
            "refined_ms": round(tr * 1e3, 3), "refined_outer": st["outer"],
            "refined_inner_total": st["inner_total"], "refined_true_resid": resr,
            "refined_inner_tol": args.inner_tol,
            "speedup_refined_vs_fp64": round(t64 / tr, 3),
            "refined_rel_err_vs_fp64": err, "refined_rel_err_vs_xtrue": errt,
            "cg_iter_bytes_fp64": metrics.cg_iter_bytes(m.n, m.nz, 3, 8),
            "cg_fp64_gbs": round(metrics.cg_iter_bytes(m.n, m.nz, 3, 8) * it64 / t64 / 1e9, 1),
        }

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
                               "kind": "port",
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

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
