#!/usr/bin/env python3
"""Benchmark of the explicit 2D Euler update on one MI355X.

Default workload (the headline configuration): modal DG P=3 (16 Legendre
modes per cell, 4x4 Gauss points), HLLC, SSP-RK3, Cockburn-Shu TVB limiter
M=50, FP64, on the four-quadrant Riemann problem of [0,1]^2 split at
(0.8, 0.8) at 1024x1024 cells (3-cell ghost ring, 1018^2 updated cells),
dt at the reference DG CFL limit nu = 0.9/(2(2P+1)).

A step is one SSP-RK3 step (three stages) over the whole grid; the state is
resident in HBM before the timed region starts.  The metric is cell-stage
updates per second (updated cells x 3 x steps / time).  The line also carries
the roofline of the dominant kernel (HIP-event durations on the library's
stream) and the CPU oracle timed on the host cores over a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "cell-stage updates/s, chosen vs plain FP64 at matched L1 err; % HBM roofline"

WORKLOADS = {
    "dg3_tvb_riemann1024": dict(method="dg", P=3, n=1024, limiter="tvb", prec="fp64"),
    "dg2_tvb_riemann1024": dict(method="dg", P=2, n=1024, limiter="tvb", prec="fp64"),
    "weno_fp64_riemann256": dict(method="weno", n=256, prec="fp64"),
    "weno_fp64_riemann1024": dict(method="weno", n=1024, prec="fp64"),
}


def algo_bytes_per_cell_stage(nm: int) -> float:
    # 4 variables x nm modes x 8 B x (stage 1: read U^n, write U^1; stages 2-3:
    # read U^n and U^(s-1), write U^(s)) = 256 nm B per cell-step over 3 stages
    return 256.0 * nm / 3.0


def build_solver(w: dict):
    import numpy as np  # noqa: F401

    from zuikakucfd_amd import meshio, zk

    m = meshio.riemann_2d(w["n"])
    if w["method"] == "dg":
        lim = {"tvb": zk.LIMITER_TVB, "persson": zk.LIMITER_PERSSON_ZS, "none": zk.LIMITER_NONE}[w["limiter"]]
        s = zk.Solver(method=zk.METHOD_DG, P=w["P"], limiter=lim, tvb_M=50.0)
        nm = (w["P"] + 1) ** 2
    else:
        prec = zk.PREC_FP64 if w["prec"] == "fp64" else zk.PREC_FP32_BULK
        s = zk.Solver(method=zk.METHOD_WENO, precision=prec)
        nm = 1
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    lam = s.max_speed()
    if w["method"] == "dg":
        cfl = 0.9 / (2.0 * (2 * w["P"] + 1))
    else:
        cfl = 0.5
    dt = cfl * m.da / lam
    return s, m, nm, dt


def cpu_baseline(w: dict, m, dt: float, budget_s: float = 20.0) -> dict:
    """Time the CPU oracle on the host cores over a bounded sample."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np

    import oracle

    threads = min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    nupd = (m.nx - 6) * (m.ny - 6)
    if w["method"] == "dg":
        lim = {"tvb": 2, "persson": 1, "none": 0}[w["limiter"]]
        o = oracle.DG.from_prims(w["P"], m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype, limiter=lim, tvb_M=50.0)
    else:
        o = oracle.FV(m.da, m.gamma, oracle.prim_to_cons(m.gamma, m.rho, m.u, m.v, m.p), m.bctype)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(dt)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 50 or el * (steps + 1) / steps > 1.5 * budget_s:
            break
    rate = steps * 3 * nupd / el
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    del np
    return {"value": rate, "unit": "cell-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} SSP-RK3 step(s) of the same {m.ny}x{m.nx} problem, C oracle with OpenMP "
                      f"({threads} threads, {model}), {el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dg3_tvb_riemann1024", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the slab decomposition is design-only); use --gpus 1.", file=sys.stderr)
        return 2

    w = WORKLOADS[args.workload]
    s, m, nm, dt = build_solver(w)
    nupd = s.updated_cells()

    s.step(dt, args.warmup)
    s.sync()
    ms = s.step_timed(dt, args.steps)
    s.sync()
    rate = args.steps * 3 * nupd / (ms / 1e3)

    # per-kernel durations with HIP events on the library's stream
    s.profile(True)
    s.step(dt, args.profile_steps)
    rep = s.profile_report()
    s.profile(False)
    dom = max(rep, key=lambda k: rep[k]["total_ms"])
    avg_ms = rep[dom]["total_ms"] / rep[dom]["launches"]
    bps = algo_bytes_per_cell_stage(nm)
    achieved = bps * nupd / (avg_ms / 1e3) / 1e9
    U = s.get_conserved()
    import numpy as np

    finite = bool(np.isfinite(U).all())
    s.close()

    line = {
        "metric": METRIC,
        "value": rate,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: four-quadrant Riemann initial data generated in-process",
        "config": {"workload": args.workload, "method": w["method"], "P": w.get("P"), "grid": [m.ny, m.nx],
                   "updated_cells": nupd, "limiter": w.get("limiter"), "precision": w["prec"], "dt": dt,
                   "parallelism": "single GPU, whole domain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                     "kernel_avg_ms": avg_ms, "algo_bytes_per_cell_stage": bps},
        "stage_hbm_frac": rate * bps / (HBM_PEAK_GBS * 1e9),
        "kernels": rep,
        "finite": finite,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, m, dt)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
