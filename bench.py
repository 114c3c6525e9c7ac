#!/usr/bin/env python3
"""Benchmark: PCISPH particle-substeps/s on one MI355X.

Workload: the 1,048,576-particle PCISPH dam break (128x64x128 simple-cubic
block at spacing 0.01, kernel radius 0.02, grid cell = kernel radius, 2x1x2
tank, dt 5e-4 fixed, TEXTBOOK predict-correct loop to <= 1% max density error,
at most 8 iterations).  One step = one substep of every particle, including
all correction iterations; value = particles * steps / wall time, inputs
already resident in HBM.

Prints one JSON line (see the keys below).  The dominant kernel's roofline is
measured live with HIP events on the solver's stream over the timed region.
The CPU baseline times the FP64 oracle (a restatement of the reference CPU
path) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M PCISPH, plain vs learned warm-start, ≤1% density err"
PEAK_HBM_GBS = 8000.0

# Compulsory FP32 bytes per particle of each kernel (xyz = 12 B, scalar = 4 B),
# excluding neighbour re-reads and the neighbour table itself.
KERNEL_BYTES = {
    "grid_hash": 36,                  # x, old cell -> new cell, key, rank, hash
    "grid_scatter": 16,               # new cell, rank, start -> slot index
    "grid_rank_move": 76,             # slot, x+id, v+p -> x+id, v+p, cell, home
    "buckets_density": 16,            # x -> rho
    "nonpressure_predict": 52,        # x, v, rho -> f, x*
    "nonpressure_features": 104,      # x, v, rho -> f, 16 features
    "predict_pressure": 28,           # x*, p -> p, rho*, q
    "pressure_force_predict": 52,     # x, q, v, f -> x*
    "pressure_force_integrate": 92,   # x, q, v, f, p -> x, v, f, f_p
    "mlp_infer": 68,                  # 16 features -> p
}


def substep_bytes(k: float) -> float:
    """Algorithmic bytes per particle-substep of the reference's PCISPH step
    with k correction iterations (FP32): 208 + 116 k."""
    return 208.0 + 116.0 * k


def run_gpu(args):
    import numpy as np

    from bubbles_amd import abi, scenes

    sc = scenes.dam_break(args.nx, args.ny, args.nz, spacing=0.01, tank=(2.0, 1.0, 2.0), seed=2)
    mode = abi.MODE_TEXTBOOK if args.mode == "textbook" else abi.MODE_REFERENCE_EXACT
    solver = abi.Solver(sc)
    warm = abi.WARM_NONE
    if args.warm == "mlp":
        from bubbles_amd import learned

        mlp = learned.load_or_train(solver, sc)
        solver.attach_mlp(mlp)
        warm = abi.WARM_MLP
    for _ in range(args.warmup):
        solver.pcisph_substep(sc.dt, args.max_it, args.tol, mode, warm)
    solver.synchronize()
    solver.profile(True)
    solver.profile_reset()
    iters, errs = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = solver.pcisph_substep(sc.dt, args.max_it, args.tol, mode, warm)
        iters.append(st.iterations)
        errs.append(st.max_density_error)
    solver.synchronize()
    t1 = time.perf_counter()
    prof = solver.profile_get()
    solver.profile(False)
    elapsed = t1 - t0
    n = sc.n
    value = n * args.steps / elapsed
    kbar = float(np.mean(iters))
    # dominant kernel
    name, rec = max(prof.items(), key=lambda kv: kv[1]["total_ms"])
    avg_ms = rec["total_ms"] / max(rec["launches"], 1)
    per_unit = KERNEL_BYTES.get(name, 0)
    achieved = per_unit * n / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": None, "avg_ms": round(avg_ms, 5), "bytes_per_particle": per_unit,
                "units_per_launch": n}
    step_bytes = substep_bytes(kbar)
    whole = {"bytes_per_particle_substep": step_bytes,
             "achieved_GBs": round(step_bytes * value / 1e9, 1),
             "frac": round(step_bytes * value / 1e9 / PEAK_HBM_GBS, 4)}
    kernels = {k: {"avg_ms": round(v["total_ms"] / max(v["launches"], 1), 5),
                   "launches": v["launches"], "total_ms": round(v["total_ms"], 3)}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["total_ms"])}
    solver.close()
    gbs, ms = abi.stream_copy(2.0 ** 30, 20)
    roofline["measured_stream_copy_GBs"] = round(gbs, 1)
    roofline["frac_of_measured"] = round(achieved / gbs, 4) if gbs > 0 else None
    return sc, value, elapsed, kbar, float(np.max(errs)), roofline, whole, kernels


def run_cpu_baseline(sc, args):
    from oracle import ref_cpu

    threads = min(16, os.cpu_count() or 1)
    sim = ref_cpu.Sim(sc, threads=threads)
    mode = ref_cpu.MODE_TEXTBOOK if args.mode == "textbook" else ref_cpu.MODE_REFERENCE_EXACT
    sim.pcisph_substep(sc.dt, args.max_it, args.tol, mode)  # warm-up
    done, t = 0, 0.0
    while done < args.cpu_steps and t < args.cpu_budget_s:
        t0 = time.perf_counter()
        sim.pcisph_substep(sc.dt, args.max_it, args.tol, mode)
        t += time.perf_counter() - t0
        done += 1
    del sim
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(sc.n * done / t, 1), "unit": "particle-steps/s", "cores": threads,
            "cpu_model": cpu_model,
            "kind": "port",
            "sample": f"{done} timed substeps (after 1 warm-up) of the same {sc.n}-particle dam "
                      f"break, FP64 oracle with linked-list cells and 100-slot buckets, "
                      f"{args.mode} loop, {threads} threads, {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mode", choices=["textbook", "exact"], default="textbook")
    ap.add_argument("--warm", choices=["none", "mlp"], default="none")
    ap.add_argument("--max-it", type=int, default=8)
    ap.add_argument("--tol", type=float, default=0.01)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-budget-s", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the PCISPH path "
              f"here has no multi-GPU mode", file=sys.stderr)
        sys.exit(2)
    sc, value, elapsed, kbar, maxerr, roofline, whole, kernels = run_gpu(args)
    cpu = None if args.no_cpu_baseline else run_cpu_baseline(sc, args)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"PCISPH dam break {sc.n} particles ({args.nx}x{args.ny}x{args.nz} "
                               f"at 0.01, h 0.02, 2x1x2 tank), dt 5e-4",
                   "n": sc.n, "mode": args.mode, "warm_start": args.warm, "max_it": args.max_it,
                   "tol": args.tol, "mean_iterations": round(kbar, 3),
                   "max_density_error": round(maxerr, 6)},
        "roofline": roofline,
        "substep_roofline": whole,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
