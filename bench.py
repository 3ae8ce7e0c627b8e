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
    "dg2_tvb_random2048": dict(method="dg", P=2, n=2048, limiter="tvb", prec="fp64", problem="random0"),
    "dg3_tvb_random2048": dict(method="dg", P=3, n=2048, limiter="tvb", prec="fp64", problem="random0"),
    "weno_fp64_riemann256": dict(method="weno", n=256, prec="fp64"),
    "weno_fp64_riemann1024": dict(method="weno", n=1024, prec="fp64"),
    "weno_fp64_jet2048x1024": dict(method="weno", problem="jet", prec="fp64"),
    "weno_fp32_jet2048x1024": dict(method="weno", problem="jet", prec="fp32"),
}


def algo_bytes_per_cell_stage(nm: int) -> float:
    # 4 variables x nm modes x 8 B x (stage 1: read U^n, write U^1; stages 2-3:
    # read U^n and U^(s-1), write U^(s)) = 256 nm B per cell-step over 3 stages
    return 256.0 * nm / 3.0


def build_solver(w: dict):
    import numpy as np  # noqa: F401

    from zuikakucfd_amd import meshio, zk

    if w.get("problem") == "jet":
        m = meshio.jet()
    elif w.get("problem") == "random0":
        m = meshio.random_riemann(0, w["n"])  # seed 0 of the 64 seeded DG P=2 problems
    else:
        m = meshio.riemann_2d(w["n"])
    if w["method"] == "dg":
        lim = {"tvb": zk.LIMITER_TVB, "persson": zk.LIMITER_PERSSON_ZS, "none": zk.LIMITER_NONE}[w["limiter"]]
        s = zk.Solver(method=zk.METHOD_DG, P=w["P"], limiter=lim, tvb_M=50.0)
        nm = (w["P"] + 1) ** 2
    else:
        prec = zk.PREC_FP64 if w["prec"] == "fp64" else zk.PREC_FP32_BULK
        # the jet needs the face positivity check (the reference scheme loses positivity on it)
        s = zk.Solver(method=zk.METHOD_WENO, precision=prec, positivity=w.get("problem") == "jet")
        nm = 1
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    lam = s.max_speed()
    if w["method"] == "dg":
        cfl = 0.9 / (2.0 * (2 * w["P"] + 1))
    else:
        cfl = 0.25 if w.get("problem") == "jet" else 0.5
    dt = cfl * m.da / lam
    return s, m, nm, dt


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` (a profiling label such as dg_fused)
    for `workload`, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (profiles/traffic.json, copy-calibrated), or None."""
    try:
        d = json.loads((ROOT / "profiles" / "traffic.json").read_text())[workload]
    except (OSError, KeyError, ValueError):
        return None
    hits = [v for k, v in d["kernels"].items() if k.startswith(kernel + "_") or k.startswith(kernel + "<")]
    if len(hits) != 1:
        return None
    return {"bytes": hits[0]["bytes"], "read_bytes": hits[0]["read_bytes"], "write_bytes": hits[0]["write_bytes"],
            "source": "profiles/" + d["source"]}


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
        o = oracle.FV(m.da, m.gamma, oracle.prim_to_cons(m.gamma, m.rho, m.u, m.v, m.p), m.bctype,
                      flux=oracle.POSITIVITY if w.get("problem") == "jet" else 0)
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


def _rel_l1(a, b) -> float:
    import numpy as np

    return float(np.abs(a - b).sum() / np.abs(b).sum())


def path_weno_riemann256(steps: int) -> dict:
    """Config 1: plain FP64 WENO5 on the shipped 256^2 four-quadrant input."""
    from zuikakucfd_amd import meshio, zk

    m = meshio.riemann_2d(256)
    s = zk.Solver(method=zk.METHOD_WENO)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    dt = 0.5 * m.da / s.max_speed()
    s.step(dt, 3)
    ms = s.step_timed(dt, steps)
    n = s.updated_cells()
    s.close()
    rate = 3 * n * steps / (ms / 1e3)
    return {"config": "WENO5 FP64, 2D-Riemann 256^2, CFL 0.5", "cell_stage_per_s": rate,
            "hbm_frac": rate * algo_bytes_per_cell_stage(1) / (HBM_PEAK_GBS * 1e9), "steps": steps}


def path_jet(steps: int) -> dict:
    """Config 3: Mach-80 jet 2048x1024, gamma 5/3, CFL 0.25; FP32-bulk WENO/HLLC
    against the plain FP64 path (both with the face positivity check, without
    which the reference scheme loses positivity on this jet), error vs FP64."""
    from zuikakucfd_amd import meshio, zk

    m = meshio.jet()
    out = {"config": "WENO5 jet 2048x1024, gamma 5/3, CFL 0.25, positivity check", "steps": steps}
    res = {}
    for name, prec in (("fp64", zk.PREC_FP64), ("fp32_bulk", zk.PREC_FP32_BULK)):
        s = zk.Solver(method=zk.METHOD_WENO, precision=prec, positivity=True)
        s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
        if name == "fp64":
            dt = 0.25 * m.da / s.max_speed()
        ms = s.step_timed(dt, steps)
        n = s.updated_cells()
        res[name] = (3 * n * steps / (ms / 1e3), s.get_conserved())
        s.close()
    out["fp64_cell_stage_per_s"] = res["fp64"][0]
    out["fp32_bulk_cell_stage_per_s"] = res["fp32_bulk"][0]
    out["speedup"] = res["fp32_bulk"][0] / res["fp64"][0]
    out["l1_rho_vs_fp64"] = _rel_l1(res["fp32_bulk"][1][0], res["fp64"][1][0])
    return out


def path_ensemble(steps: int, ninst: int = 256) -> dict:
    """Config 5: 256 seeded four-quadrant problems, N uniform over
    {128,192,256,384,512}, per-instance CFL 0.5 dt; FP32-bulk batched ensemble
    (persistent launch, graph replay, autotuned tiles) against the plain FP64
    per-instance loop; density error vs the FP64 loop."""
    import numpy as np

    from zuikakucfd_amd import autotune, meshio, zk

    rng = np.random.default_rng(2024)
    sizes = rng.choice([128, 192, 256, 384, 512], size=ninst)
    ms = [meshio.random_riemann(1000 + k, int(n)) for k, n in enumerate(sizes)]

    def make(prec):
        E = zk.Ensemble(precision=prec, positivity=True)
        for m in ms:
            E.add(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype, cfl=0.5)
        E.finalize()
        return E

    tile, tune_ms = autotune.tune_ensemble(lambda: make(zk.PREC_FP32_BULK),
                                           key=f"ens-fp32-{ninst}-{int(sum(int(n) ** 2 for n in sizes))}")
    E = make(zk.PREC_FP32_BULK)
    E.set_tuning(tile=tile, use_graph=True)
    _, ncell = E.size()
    E.step(1)
    E.sync()
    E2 = make(zk.PREC_FP32_BULK)
    E2.set_tuning(tile=tile, use_graph=True)
    t0 = time.perf_counter()
    E2.step(steps)
    E2.sync()
    t_ens = time.perf_counter() - t0
    del E
    # plain FP64 per-instance loop: one handle per problem, on-device CFL stepping
    solvers = []
    for m in ms:
        s = zk.Solver(method=zk.METHOD_WENO, positivity=True)
        s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
        solvers.append(s)
    t0 = time.perf_counter()
    for s in solvers:
        s.step_cfl(0.5, 1e30, steps)
    t_loop = time.perf_counter() - t0
    # Some draws form a vacuum (a strong rarefaction between diverging
    # quadrants) within the run; the reference scheme, face positivity check
    # included, is not positivity-preserving there and goes non-finite in
    # FP64 as well.  Errors are taken over instances finite in both runs.
    errs = []
    nonfinite, nonfinite64 = 0, 0
    for k in range(ninst):
        U, _ = E2.get(k)
        U64 = solvers[k].get_conserved()
        bad32, bad64 = not np.isfinite(U).all(), not np.isfinite(U64).all()
        nonfinite += bad32
        nonfinite64 += bad64
        if not (bad32 or bad64):
            errs.append(_rel_l1(U[0], U64[0]))
    for s in solvers:
        s.close()
    E2.close()
    r_ens = 3 * ncell * steps / t_ens
    r_loop = 3 * ncell * steps / t_loop
    return {"config": f"{ninst} WENO5 four-quadrant problems, N in {{128..512}}, {ncell} updated cells, CFL 0.5",
            "steps": steps, "fp32_batched_cell_stage_per_s": r_ens, "fp64_loop_cell_stage_per_s": r_loop,
            "speedup": r_ens / r_loop, "l1_rho_vs_fp64_max": max(errs),
            "l1_rho_vs_fp64_median": float(np.median(errs)), "l1_rho_vs_fp64_p95": float(np.percentile(errs, 95)),
            "instances_above_1e-4": int(sum(e > 1e-4 for e in errs)), "nonfinite_instances": nonfinite,
            "nonfinite_instances_fp64_loop": nonfinite64,
            "positivity_check": True, "tile": tile, "autotune_ms": tune_ms,
            "timing": "wall clock around the launches with a device sync at each end"}


def path_learned(steps: int, problems: int, n: int = 2048) -> dict:
    """Config 4: DG P=2 on seeded random four-quadrant problems at n^2; the
    learned troubled-cell indicator (shipped network, band [0.2, 0.8], 2%
    audit) against the plain FP64 TVB limiter from the same state with the
    same dt: rates, density L1 difference, the network's miss and false-flag
    rates on audited cells."""
    from zuikakucfd_amd import learned

    if not learned.WEIGHTS.exists():
        return {"error": f"no trained network at {learned.WEIGHTS.name}"}
    mlp = learned.load()
    rows = [learned.evaluate(mlp, seed, n, steps) for seed in range(problems)]
    mlp.close()
    tvb = sum(r["tvb_cell_stage_per_s"] for r in rows) / len(rows)
    lrn = sum(r["learned_cell_stage_per_s"] for r in rows) / len(rows)
    tot = {k: sum(r[k] for r in rows) for k in ("audited", "agree", "miss", "false_flag", "net_flagged", "in_band",
                                                 "candidates", "audited_positive")}
    return {"config": f"DG P=2 {n}^2, TVB M=50 vs learned indicator (band [0.2, 0.8], 2% audit), "
                      f"seeds 0..{problems - 1} of 64, {steps} steps each",
            "steps": steps, "problems": problems, "tvb_cell_stage_per_s": tvb, "learned_cell_stage_per_s": lrn,
            "speedup": lrn / tvb, "l1_rho_vs_tvb_max": max(r["l1_rho_vs_tvb"] for r in rows),
            **learned.rates(tot),
            "candidate_fraction": sum(r["candidate_fraction"] for r in rows) / len(rows),
            "band_fraction": tot["in_band"] / max(tot["candidates"], 1),
            "finite": all(r["finite"] for r in rows)}


def device_copy(nbytes: int) -> int:
    """One device-to-device hipMemcpy of nbytes (the runtime's plain copy
    kernel), used to scale the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dg3_tvb_riemann1024", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-paths", action="store_true", help="skip the other configurations")
    ap.add_argument("--path-steps", type=int, default=20)
    ap.add_argument("--ensemble-steps", type=int, default=200,
                    help="steps of every instance in the batched-ensemble configuration (200 as specified)")
    ap.add_argument("--learned-problems", type=int, default=4,
                    help="seeded problems of the DG P=2 learned-indicator configuration (of 64)")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB on the device (a plain copy "
                         "kernel of known size, to scale FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this library has no multi-GPU path "
              "(the slab decomposition is design-only); use --gpus 1.", file=sys.stderr)
        return 2

    w = WORKLOADS[args.workload]
    s, m, nm, dt = build_solver(w)
    nupd = s.updated_cells()

    import numpy as np

    # Every timed block starts again from the initial data, so all blocks time
    # the same window of the solution (the count of limited cells grows with
    # time) and none can run into a state the scheme has lost.
    def restart(first: bool = False) -> None:
        if not first:
            s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
        s.step(dt, args.warmup)
        s.sync()

    finite_blocks = []
    restart(first=True)
    ms = s.step_timed(dt, args.steps)
    s.sync()
    finite_blocks.append(bool(np.isfinite(s.get_conserved()).all()))
    rate = args.steps * 3 * nupd / (ms / 1e3)
    # five more timed blocks of the same length: their median beside `value`
    reps = []
    for _ in range(5):
        restart()
        reps.append(s.step_timed(dt, args.steps))
        finite_blocks.append(bool(np.isfinite(s.get_conserved()).all()))
    reps.sort()
    rate_median5 = args.steps * 3 * nupd / (reps[2] / 1e3)

    # per-kernel durations with HIP events on the library's stream
    restart()
    s.profile(True)
    s.step(dt, args.profile_steps)
    rep = s.profile_report()
    s.profile(False)
    dom = max(rep, key=lambda k: rep[k]["total_ms"])
    avg_ms = rep[dom]["total_ms"] / rep[dom]["launches"]
    bps = algo_bytes_per_cell_stage(nm)
    achieved = bps * nupd / (avg_ms / 1e3) / 1e9
    U = s.get_conserved()
    finite_blocks.append(bool(np.isfinite(U).all()))
    finite = all(finite_blocks)
    # The C ABI takes and returns host buffers.  The rate with the transfers
    # of the executables' cadence (the conserved planes read back every 100
    # steps) is reported beside `value`, which has the state resident in HBM.
    import time

    get_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        s.get_conserved()
        get_ms.append((time.perf_counter() - t0) * 1e3)
    t0 = time.perf_counter()
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    s.sync()
    load_ms = (time.perf_counter() - t0) * 1e3
    s.close()
    boundary = {"load_ms": load_ms, "get_conserved_ms": min(get_ms),
                "cell_stage_per_s_with_output_every_100_steps":
                    100 * 3 * nupd / ((100 * ms / args.steps + min(get_ms)) / 1e3),
                "note": "host buffers through the C ABI (zk_load, zk_get_conserved); wall clock"}

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
        "dtype": "f64" if w["prec"] == "fp64" else "f32 bulk, f64 state",
        "data": "synthetic: " + {"jet": "Mach-80 jet inflow (2048x1024, gamma 5/3)",
                                 "random0": "seeded random four-quadrant Riemann problem (seed 0)"}.get(
                                     w.get("problem"), "four-quadrant Riemann initial data") + " generated in-process",
        "config": {"workload": args.workload, "method": w["method"], "P": w.get("P"), "grid": [m.ny, m.nx],
                   "updated_cells": nupd, "limiter": w.get("limiter"), "precision": w["prec"], "dt": dt,
                   "parallelism": "single GPU, whole domain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                     "kernel_avg_ms": avg_ms, "algo_bytes_per_cell_stage": bps,
                     "algo_bytes_per_launch": bps * nupd},
        "stage_hbm_frac": rate * bps / (HBM_PEAK_GBS * 1e9),
        "median_of_5": {"value": rate_median5, "ms_per_step": reps[2] / args.steps,
                        "note": "five further timed blocks of the same step count (HIP events), each "
                                "restarted from the initial data after the same warm-up"},
        "kernels": rep,
        "finite": finite,
        "last_step": args.warmup + max(args.steps, args.profile_steps),
        "boundary": boundary,
    }
    tr = measured_traffic(args.workload, dom)
    if tr:
        line["roofline"]["traffic"] = tr["bytes"]
        line["roofline"]["traffic_detail"] = tr
    if not args.no_paths:
        line["paths"] = {"weno_fp64_riemann256": path_weno_riemann256(4 * args.path_steps),
                         "jet_fp32_bulk": path_jet(args.path_steps),
                         "ensemble256": path_ensemble(args.ensemble_steps),
                         "dg2_learned2048": path_learned(args.path_steps, args.learned_problems)}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, m, dt)
    if args.calibrate_copy_mib > 0:
        line["calibration_copy_bytes"] = device_copy(args.calibrate_copy_mib << 20)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
