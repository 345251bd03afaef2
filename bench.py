#!/usr/bin/env python3
"""Particle-steps/s of the PBF hot path on one MI355X.

Workload: the 1,048,576-particle PBF dam break (128 x 128 x 64 jittered block at spacing h/2, h = 0.1, in a
12.8 x 6.4 x 6.4 box; 4 Jacobi iterations, eps = 600, s_corr k = 0.1 n = 4, XSPH c = 0.01, dt = 1/120).
A step is one pass of the hot path over all particles: prediction, grid rebuild, neighbour list, Jacobi
iterations, velocity update and XSPH.  Inputs are resident in HBM before the timed region starts.

Prints one JSON line.  `roofline` is measured with HIP events around the dominant kernel class during the
timed steps; `cpu_baseline` times the oracle (single-threaded FP32 restatement of the reference CPU path)
on one step of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per launch: bytes per particle and per grid cell (each array streamed once).
ALG_BYTES = {
    "predict": (116, 16),      # density (12 + 4) + viscosity (12 + 12 + 4 + 12) + integrate (36 + 24)
    "grid": (96, 24),          # hash + stable counting sort + reorder
    "list_lambda": (20, 8),    # lambda pass
    "lambda": (20, 8),
    "dp": (28, 8),             # position correction + apply + collide
    "velocity": (36, 0),
    "xsph": (40, 8),
}


def fused_grid_cells(sp, cell: float) -> int:
    g = 1
    for k in range(3):
        ws = np.float32(sp.WorldMax[k]) - np.float32(sp.WorldOrigin[k])
        g *= int(ws / np.float32(cell)) + 1
    return g


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="pbf_1m")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU path is a design note (no figure for N > 1)",
              file=sys.stderr)
        return 2

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from fltk_sph_ice_amd import _abi, scene
    from fltk_sph_ice_amd.solvers import PBFSolver

    sc = scene.pbf_1m() if args.workload == "pbf_1m" else scene.pbf_dam_break()
    pos, vel = sc.particles()
    n = sc.n
    sp = sc.params()
    solver = PBFSolver(sc, pos, vel)
    solver.step(args.warmup)
    solver.sync()
    solver.option(_abi.RX_PBF_OPT_PROFILE, 1)
    solver.profile()  # reset
    solver.sync()
    t0 = time.perf_counter()
    solver.step(args.steps)
    solver.sync()
    t1 = time.perf_counter()
    prof = solver.profile()
    stats = solver.stats()
    solver.option(_abi.RX_PBF_OPT_PROFILE, 0)
    # judge the last step: one more prediction pass measures the density of the final positions
    solver.step(1)
    judge = solver.stats()
    elapsed = t1 - t0
    value = n * args.steps / elapsed

    G = fused_grid_cells(sp, 1.1 * sc.h)
    per_kernel = {}
    for name, (ms, cnt) in prof.items():
        bp, bc = ALG_BYTES.get(name, (0, 0))
        avg_ms = ms / max(cnt, 1)
        alg = bp * n + bc * G
        per_kernel[name] = {"avg_ms": round(avg_ms, 5), "launches": cnt,
                            "alg_GBs": round(alg / (avg_ms * 1e-3) / 1e9, 1) if avg_ms > 0 else None}
    dom = max(prof.items(), key=lambda kv: kv[1][0])[0] if prof else None
    roof = None
    if dom:
        bp, bc = ALG_BYTES[dom]
        avg_s = prof[dom][0] / prof[dom][1] * 1e-3
        ach = (bp * n + bc * G) / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                "alg_bytes_per_launch": bp * n + bc * G}

    cpu = None
    if not args.no_cpu_baseline:
        import ctypes as C

        import oracle

        P = oracle.params_from(sp, sc)
        tc = time.perf_counter()
        oracle.pbf_step(pos, vel, P)
        tcpu = time.perf_counter() - tc
        cpu = {"value": round(n / tcpu, 1), "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"1 step of the same {n}-particle workload ({tcpu:.1f} s), oracle/liboracle.so FP32, 1 thread"}

    out = {
        "metric": "particle-steps/s at <=1% density error, plain vs learned-warm-start PBF",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{sc.name}: {sc.block[0]}x{sc.block[1]}x{sc.block[2]} PBF dam break, {n} particles, "
                               f"h={sc.h}, {sc.iterations} Jacobi iterations, eps={sc.epsilon}, dt=1/120",
                   "solver": "plain"},
        "density_error": {"max_start_of_last_step": float(stats[0]), "max_after_last_step": float(judge[0]),
                          "mean_abs_after_last_step": float(judge[23]),
                          "per_iteration_last_step": [float(x) for x in stats[1:1 + sc.iterations]],
                          "within_1pct": bool(judge[0] <= 0.01)},
        "roofline": roof,
        "kernels": per_kernel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
