#!/usr/bin/env python3
"""Benchmark of the explicit 2D Euler update on one MI355X.

Default workload (the headline configuration): modal DG P=3 (16 Legendre
modes per cell, 4x4 Gauss points), HLLC, SSP-RK3, Cockburn-Shu TVB limiter
M=50, FP64, on the four-quadrant Riemann problem of [0,1]^2 split at
(0.8, 0.8) at 1024x1024 cells (3-cell ghost ring, 1018^2 updated cells),
dt at the reference DG CFL limit nu = 0.9/(2(2P+1)).

A step is one SSP-RK3 step (three stages) over the whole grid; the state is
resident in HBM before the timed region starts.  The metric is cell-stage
updates per second (updated cells x 3 x steps / time).  The timed window sits
in the middle of the run: W untimed steps (default: half of the steps to the
workload's end time) advance the solution, then K steps are timed, because the
cost per step grows as limiting spreads; the first steps of the run are timed
too and reported beside it.  The line also carries the roofline of the
dominant kernel (HIP-event durations on the library's stream) and the CPU
oracle timed on the host cores over a bounded sample (in a separate process,
while the GPU advances to the timed window).

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
FP64_PEAK_TFS = 78.6  # MI355X dense FP64 (vector and matrix alike)
METRIC = "cell-stage updates/s, chosen vs plain FP64 at matched L1 err; % HBM roofline"

# t_end: the end time of the configuration (the default timed window sits at
# half of it); end_steps instead where the configuration fixes a step count
WORKLOADS = {
    "dg3_tvb_riemann1024": dict(method="dg", P=3, n=1024, limiter="tvb", prec="fp64", t_end=0.8),
    "dg2_tvb_riemann1024": dict(method="dg", P=2, n=1024, limiter="tvb", prec="fp64", t_end=0.8),
    "dg2_tvb_random2048": dict(method="dg", P=2, n=2048, limiter="tvb", prec="fp64", problem="random0",
                               end_steps=2000),
    "dg3_tvb_random2048": dict(method="dg", P=3, n=2048, limiter="tvb", prec="fp64", problem="random0",
                               end_steps=2000),
    "weno_fp64_riemann256": dict(method="weno", n=256, prec="fp64", t_end=0.8),
    "weno_fp64_riemann1024": dict(method="weno", n=1024, prec="fp64", t_end=0.8),
    "weno_fp64_jet2048x1024": dict(method="weno", problem="jet", prec="fp64", t_end=0.07),
    "weno_fp32_jet2048x1024": dict(method="weno", problem="jet", prec="fp32", t_end=0.07),
}


def algo_bytes_per_cell_stage(nm: int) -> float:
    # 4 variables x nm modes x 8 B x (stage 1: read U^n, write U^1; stages 2-3:
    # read U^n and U^(s-1), write U^(s)) = 256 nm B per cell-step over 3 stages
    return 256.0 * nm / 3.0


def build_mesh(w: dict):
    from zuikakucfd_amd import meshio

    if w.get("problem") == "jet":
        return meshio.jet()
    if w.get("problem") == "random0":
        return meshio.random_riemann(0, w["n"])  # seed 0 of the 64 seeded DG P=2 problems
    return meshio.riemann_2d(w["n"])


def build_solver(w: dict):
    from zuikakucfd_amd import zk

    m = build_mesh(w)
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
    if not hits:
        return None
    # several instances of one kernel (e.g. the first RK stage and the later
    # ones): the average per launch, weighted by their launch counts
    wsum = sum(h.get("dispatches", 1) for h in hits)
    avg = {key: sum(h[key] * h.get("dispatches", 1) for h in hits) / wsum
           for key in ("bytes", "read_bytes", "write_bytes")}
    avg["instances"] = len(hits)
    avg["source"] = "profiles/" + d["source"]
    return avg


def measured_flops(workload: str, kernel: str):
    """Executed FP64 operations per launch of `kernel` for `workload` (VALU and
    MFMA, from the committed rocprofv3 counter summary profiles/flops.json),
    weighted over the kernel's instances by their launch counts, or None."""
    try:
        d = json.loads((ROOT / "profiles" / "flops.json").read_text())[workload]
    except (OSError, KeyError, ValueError):
        return None
    hits = [v for k, v in d["kernels"].items() if k.startswith(kernel + "_") or k.startswith(kernel + "<")]
    if not hits:
        return None
    wsum = sum(h.get("dispatches", 1) for h in hits)
    avg = {key: sum(h[key] * h.get("dispatches", 1) for h in hits) / wsum for key in ("flops", "valu_flops", "mfma_flops")}
    avg["source"] = "profiles/" + d["source"]
    return avg


def _cpu_model() -> str:
    """The first processor's model name as /proc/cpuinfo reports it (a
    virtual machine may shorten it), with its family/model numbers and widest
    vector extension, which identify the part when the name does not."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if not line.strip():
                break
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except OSError:
        return "unknown"
    flags = info.get("flags", "").split()
    simd = "avx512f" if "avx512f" in flags else "avx2" if "avx2" in flags else "sse"
    return (f"{info.get('model name', 'unknown')} (family {info.get('cpu family', '?')}, "
            f"model {info.get('model', '?')}, {simd})")


def _native_oracle() -> tuple[str | None, str]:
    """zk_oracle.c compiled on this host with the reference Makefile's flags
    (Makefile:7, -O3 -march=native -mtune=native -funroll-loops) plus
    -fopenmp, into a temporary folder; (path, flags) or (None, reason)."""
    import shutil
    import subprocess
    import tempfile

    flags = "-O3 -march=native -mtune=native -funroll-loops -fopenmp"
    cc = shutil.which("gcc")
    if not cc:
        return None, "no gcc on this host"
    out = Path(tempfile.mkdtemp(prefix="zk_oracle_native_")) / "liboracle_native.so"
    r = subprocess.run([cc, *flags.split(), "-fPIC", "-std=gnu11", "-shared", "-o", str(out),
                        str(ROOT / "oracle" / "zk_oracle.c"), "-lm"], capture_output=True, text=True)
    if r.returncode != 0:
        return None, "native build failed: " + r.stderr[-200:]
    return str(out), flags


def cpu_baseline(w: dict, m, dt: float, warmup: int = 5, steps: int = 20, max_s: float = 90.0) -> dict:
    """Time the CPU oracle on the host cores: `warmup` untimed steps, then
    `steps` timed steps of the same problem (fewer if they would take longer
    than max_s, never fewer than 10), built with the reference's flags."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    lib, flags = _native_oracle()
    if lib:
        oracle.use_library(lib)
        build = flags
    else:
        build = "oracle/liboracle.so (-O3 -march=x86-64-v3 -funroll-loops -fopenmp; " + flags + ")"
    threads = int(os.environ.get("OMP_NUM_THREADS") or (os.cpu_count() or 1))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    nupd = (m.nx - 6) * (m.ny - 6)
    if w["method"] == "dg":
        lim = {"tvb": 2, "persson": 1, "none": 0}[w["limiter"]]
        o = oracle.DG.from_prims(w["P"], m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype, limiter=lim, tvb_M=50.0)
    else:
        o = oracle.FV(m.da, m.gamma, oracle.prim_to_cons(m.gamma, m.rho, m.u, m.v, m.p), m.bctype,
                      flux=oracle.POSITIVITY if w.get("problem") == "jet" else 0)
    t0 = time.perf_counter()
    o.step(dt, warmup)
    per_step = (time.perf_counter() - t0) / max(warmup, 1)
    n = steps if per_step * steps <= max_s else max(10, int(max_s / per_step))
    t0 = time.perf_counter()
    o.step(dt, n)
    el = time.perf_counter() - t0
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"value": n * 3 * nupd / el, "unit": "cell-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} timed SSP-RK3 steps after {warmup} untimed of the same {m.ny}x{m.nx} problem and dt, "
                      f"the C oracle (OpenMP, {threads} threads) built {build}; {el:.1f} s",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": aff,
            "threads_note": "OMP_NUM_THREADS as this host sets it (a host that shares its cores among "
                            "jobs sets it to the job's share); nproc as os.cpu_count() reports it"}


def _cpu_baseline_child(args) -> int:
    """Entry of the CPU-baseline process (no GPU use): prints one JSON line."""
    w = WORKLOADS[args.workload]
    m = build_mesh(w)
    print(json.dumps(cpu_baseline(w, m, args.cpu_dt)), flush=True)
    return 0


def start_cpu_baseline(workload: str, dt: float):
    import subprocess

    return subprocess.Popen([sys.executable, str(Path(__file__).resolve()), "--cpu-baseline-only",
                             "--workload", workload, "--cpu-dt", repr(dt)],
                            stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def finish_cpu_baseline(proc) -> dict:
    out, err = proc.communicate()
    if proc.returncode != 0:
        return {"error": f"CPU baseline process failed ({proc.returncode}): {err[-300:]}"}
    return json.loads(out.strip().splitlines()[-1])


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
    out["fp64_hbm_frac"] = res["fp64"][0] * algo_bytes_per_cell_stage(1) / (HBM_PEAK_GBS * 1e9)
    fl = measured_flops("weno_fp64_jet2048x1024", "fv_stage")
    if fl:
        # executed FP64 operations per cell-stage (counted) at the FP64 path's rate
        per_cell = fl["flops"] / n
        out["fp64_flops_per_cell_stage"] = per_cell
        out["fp64_fp_frac"] = res["fp64"][0] * per_cell / (FP64_PEAK_TFS * 1e12)
    return out


def path_ensemble(steps: int, ninst: int = 256) -> dict:
    """Config 5: 256 seeded four-quadrant problems, N uniform over
    {128,192,256,384,512}, per-instance CFL 0.5 dt, 200 steps.  Three rates
    separate the two gains: the plain FP64 per-instance loop (one handle per
    problem, zk_step_cfl with its CFL step captured beforehand and replayed,
    a host read of the step scalars every 16 steps), the FP64
    batched ensemble (batching alone) and the FP32-bulk batched ensemble
    (batching and precision; persistent launch, graph replay, autotuned
    tiles).  Density error of FP32 bulk against the FP64 loop per instance."""
    import numpy as np

    from zuikakucfd_amd import autotune, meshio, zk

    rng = np.random.default_rng(2024)
    sizes = rng.choice([128, 192, 256, 384, 512], size=ninst)
    ms = [meshio.random_riemann(1000 + k, int(n)) for k, n in enumerate(sizes)]

    def make(prec, tile=None):
        E = zk.Ensemble(precision=prec, positivity=True)
        for m in ms:
            E.add(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype, cfl=0.5)
        E.finalize()
        if tile:
            E.set_tuning(tile=tile, use_graph=True)
        return E

    def timed(prec, tile):
        W = make(prec, tile)   # loads the kernels and runs them once
        W.step(1)
        W.sync()
        W.close()
        E = make(prec, tile)
        E.sync()
        t0 = time.perf_counter()
        E.step(steps)          # includes capturing the step graph (about a millisecond)
        E.sync()
        return E, time.perf_counter() - t0

    tile, tune_ms = autotune.tune_ensemble(lambda: make(zk.PREC_FP32_BULK),
                                           key=f"ens-fp32-{ninst}-{int(sum(int(n) ** 2 for n in sizes))}")
    tile64, _ = autotune.tune_ensemble(lambda: make(zk.PREC_FP64),
                                       key=f"ens-fp64-{ninst}-{int(sum(int(n) ** 2 for n in sizes))}")
    E32, t_ens = timed(zk.PREC_FP32_BULK, tile)
    _, ncell = E32.size()
    E64, t_ens64 = timed(zk.PREC_FP64, tile64)
    E64.close()
    # plain FP64 per-instance loop: one handle per problem, on-device CFL stepping
    solvers = []
    for m in ms:
        s = zk.Solver(method=zk.METHOD_WENO, positivity=True)
        s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
        solvers.append(s)
    # every handle captures its CFL step outside the timing (one step, then
    # the initial state back: zk_set_state keeps the buffers and the graph)
    for s in solvers:
        q0 = s.get_state()
        s.step_cfl(0.5, 1e30, 1)
        s.set_state(q0)
    for s in solvers:
        s.sync()
    t0 = time.perf_counter()
    for s in solvers:
        s.step_cfl(0.5, 1e30, steps)
    t_loop = time.perf_counter() - t0
    errs = []
    nonfinite, nonfinite64 = 0, 0
    for k in range(ninst):
        U, _ = E32.get(k)
        U64 = solvers[k].get_conserved()
        bad32, bad64 = not np.isfinite(U).all(), not np.isfinite(U64).all()
        nonfinite += bad32
        nonfinite64 += bad64
        if not (bad32 or bad64):
            errs.append(_rel_l1(U[0], U64[0]))
    for s in solvers:
        s.close()
    E32.close()
    r_ens = 3 * ncell * steps / t_ens
    r_ens64 = 3 * ncell * steps / t_ens64
    r_loop = 3 * ncell * steps / t_loop
    above = int(sum(e > 1e-4 for e in errs))
    return {"config": f"{ninst} WENO5 four-quadrant problems, N in {{128..512}}, {ncell} updated cells, CFL 0.5",
            "steps": steps, "fp32_batched_cell_stage_per_s": r_ens, "fp64_batched_cell_stage_per_s": r_ens64,
            "fp64_loop_cell_stage_per_s": r_loop,
            "speedup": r_ens / r_loop, "batching_gain": r_ens64 / r_loop, "precision_gain": r_ens / r_ens64,
            "l1_rho_vs_fp64_max": max(errs), "l1_rho_vs_fp64_median": float(np.median(errs)),
            "l1_rho_vs_fp64_p95": float(np.percentile(errs, 95)), "instances_above_1e-4": above,
            "matched": above == 0,
            "matched_note": "matched = every instance's density L1 against the FP64 loop within 1e-4; draw 101 "
                            "of C5 is the exception, whose FP64 run moves by 4.3e-4 under a 1e-7 perturbation "
                            "of its initial state",
            "nonfinite_instances": nonfinite, "nonfinite_instances_fp64_loop": nonfinite64,
            "positivity_check": True, "tile": tile, "tile_fp64": tile64, "autotune_ms": tune_ms,
            "timing": "wall clock around the launches with a device sync at each end, after a warm-up run "
                      "of the same kernels"}


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
            "finite": all(r["finite"] for r in rows),
            "sweep_64x2000": recorded_c4_sweep()}


def recorded_c4_sweep() -> dict | None:
    """The summary of the recorded full C4 sweep (64 problems x 2,000 steps,
    `python -m zuikakucfd_amd.learned --c4 0 64`, about half an hour of GPU
    time, too long for this bench): finite counts, rates and the network's
    rates over all audited cells."""
    p = ROOT / "zuikakucfd_amd" / "data" / "c4_sweep_2048_2000.json"
    try:
        d = json.loads(p.read_text())["summary"]
    except (OSError, KeyError, ValueError):
        return None
    d["recorded"] = f"zuikakucfd_amd/data/{p.name} (not measured by this run)"
    return d


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


def stream_reference(state_bytes: int) -> dict:
    """Streaming bandwidth this GPU gives for the stage's read/write mixes, over
    arrays the size of one state array (zk_stream_bandwidth): a copy (1 read +
    1 write: the first RK stage reads U^n and writes U^1) and an add (2 reads +
    1 write: the later stages).  `stage_mix_gbs` is the rate at which those
    kernels would move one launch's algorithmic bytes (the three stages'
    average): the practical ceiling of a stage kernel that moves exactly those
    bytes."""
    from zuikakucfd_amd import zk

    copy, add = zk.stream_bandwidth(state_bytes)
    t = (2 * state_bytes / copy + 2 * 3 * state_bytes / add) / 3 / 1e9  # seconds per launch, stage average
    return {"copy_gbs": copy, "add_gbs": add, "stage_mix_gbs": (8 * state_bytes / 3) / t / 1e9,
            "array_bytes": state_bytes,
            "note": "zk_stream_bandwidth: a copy (1 read + 1 write) and an add (2 reads + 1 write) over arrays of one "
                    "state's size, 16-byte accesses, best of 5 after a first run; stage_mix_gbs weights them as the "
                    "3 RK stages move their bytes"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="steps of each timed block")
    ap.add_argument("--warmup", type=int, default=-1,
                    help="untimed steps before the timed window (default: half of the steps to the workload's "
                         "end time, or of its step count)")
    ap.add_argument("--early-warmup", type=int, default=5,
                    help="untimed steps before the early window (the first steps of the run)")
    ap.add_argument("--workload", default="dg3_tvb_riemann1024", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-paths", action="store_true", help="skip the other configurations")
    ap.add_argument("--path-steps", type=int, default=20)
    ap.add_argument("--ensemble-steps", type=int, default=200,
                    help="steps of every instance in the batched-ensemble configuration (200 as specified)")
    ap.add_argument("--learned-problems", type=int, default=4,
                    help="seeded problems of the DG P=2 learned-indicator configuration (of 64)")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--save-state", default=None, metavar="NPY",
                    help="write the state reached at the timed window (the warm-up's end) to NPY")
    ap.add_argument("--resume-state", default=None, metavar="NPY",
                    help="start the timed window from a state written by --save-state for the same workload and "
                         "warm-up, instead of advancing to it (short runs under a profiler)")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the streaming-bandwidth reference measured beside the roofline")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB on the device (a plain copy "
                         "kernel of known size, to scale FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--cpu-baseline-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-dt", type=float, default=0.0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_only:
        return _cpu_baseline_child(args)
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this library has no multi-GPU path "
              "(the slab decomposition is design-only); use --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np

    w = WORKLOADS[args.workload]
    s, m, nm, dt = build_solver(w)
    nupd = s.updated_cells()
    if args.warmup < 0:
        args.warmup = int(0.5 * w["t_end"] / dt) if "t_end" in w else w["end_steps"] // 2
    # the CPU oracle runs in its own process while the GPU advances
    cpu_proc = None if args.no_cpu else start_cpu_baseline(args.workload, dt)

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    # early window: the first steps of the run
    s.step(dt, args.early_warmup)
    s.sync()
    early_ms = s.step_timed(dt, args.steps)
    early = {"value": args.steps * 3 * nupd / (early_ms / 1e3), "ms_per_step": early_ms / args.steps,
             "steps": f"{args.early_warmup + 1}..{args.early_warmup + args.steps}"}
    # advance, untimed, to the timed window
    done = args.early_warmup + args.steps
    t_adv = time.perf_counter()
    if args.resume_state:
        s.set_state(np.load(args.resume_state))
        done = max(done, args.warmup)
        log(f"resumed at step {done} from {args.resume_state}")
    while done < args.warmup:
        k = min(1000, args.warmup - done)
        s.step(dt, k)
        s.sync()
        done += k
        if done % 4000 < 1000 or done == args.warmup:
            log(f"advanced to step {done} of {args.warmup} ({time.perf_counter() - t_adv:.1f} s)")
    advance_s = time.perf_counter() - t_adv
    if args.save_state:
        np.save(args.save_state, s.get_state())
    cpu = finish_cpu_baseline(cpu_proc) if cpu_proc else None  # the timed window gets the host to itself
    ms = s.step_timed(dt, args.steps)
    rate = args.steps * 3 * nupd / (ms / 1e3)
    # four more timed blocks right after it: the median of the five beside `value`
    reps = [ms] + [s.step_timed(dt, args.steps) for _ in range(4)]
    reps.sort()
    rate_median5 = args.steps * 3 * nupd / (reps[2] / 1e3)
    # per-kernel durations with HIP events on the library's stream, at the same point of the run
    s.profile(True)
    s.step(dt, args.profile_steps)
    rep = s.profile_report()
    s.profile(False)
    last_step = max(done, args.warmup) + 5 * args.steps + args.profile_steps
    dom = max(rep, key=lambda k: rep[k]["total_ms"])
    avg_ms = rep[dom]["total_ms"] / rep[dom]["launches"]
    bps = algo_bytes_per_cell_stage(nm)
    achieved = bps * nupd / (avg_ms / 1e3) / 1e9
    U = s.get_conserved()
    finite = bool(np.isfinite(U).all())
    # The C ABI takes and returns host buffers.  The rate with the transfers
    # of the executables' cadence (the conserved planes read back every 100
    # steps) is reported beside `value`, which has the state resident in HBM.
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
    stream = None if args.no_stream else stream_reference(4 * nm * m.ny * m.nx * 8)
    boundary = {"load_ms": load_ms, "get_conserved_ms": min(get_ms),
                "cell_stage_per_s_with_output_every_100_steps":
                    100 * 3 * nupd / ((100 * ms / args.steps + min(get_ms)) / 1e3),
                "note": "host buffers through the C ABI (zk_load, zk_get_conserved); wall clock"}

    lim_note = None
    if w["method"] == "dg" and w.get("limiter") == "tvb":
        lim_note = ("TVB M=50 on the linear modes" + (", with the face-average test at P=3" if w["P"] >= 3 else "")
                    + ", then Zhang-Shu positivity")
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
                   "updated_cells": nupd, "limiter": w.get("limiter"), "limiter_rule": lim_note,
                   "precision": w["prec"], "dt": dt, "timed_steps": f"{args.warmup + 1}..{args.warmup + args.steps}",
                   "parallelism": "single GPU, whole domain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                     "kernel_avg_ms": avg_ms, "algo_bytes_per_cell_stage": bps,
                     "algo_bytes_per_launch": bps * nupd},
        "stage_hbm_frac": rate * bps / (HBM_PEAK_GBS * 1e9),
        "stream_reference": stream,
        "median_of_5": {"value": rate_median5, "ms_per_step": reps[2] / args.steps,
                        "note": "the timed block and four more right after it (HIP events)"},
        "early_window": early,
        "advance_s": advance_s,
        "kernels": rep,
        "finite": finite,
        "last_step": last_step,
        "boundary": boundary,
    }
    if stream:
        # the same bytes moved by plain streaming kernels: the practical ceiling
        line["roofline"]["frac_of_stream_mix"] = achieved / stream["stage_mix_gbs"]
    if args.workload.startswith("dg3_tvb"):
        # what the counters of this kernel show (committed rocprofv3 summaries)
        line["roofline"]["measured_limit"] = (
            "memory access pattern and FP64 issue together, not HBM bandwidth: with every global access kept in "
            "the L2 the same instruction stream runs at about 306 us per launch against about 376 measured; the "
            "waves issue 33-37% of their cycles, wait to issue 38-42% and wait on memory or barriers 22-30% "
            "(profiles/dg3_tvb_riemann1024_stall_pmc.txt)")
    if args.resume_state:
        line["resumed"] = ("the timed window started from the state a --save-state run reached at step "
                           f"{args.warmup} (the untimed advance skipped)")
    fl = measured_flops(args.workload, dom)
    if fl and w["prec"] == "fp64":
        # the second fraction: executed FP64 operations against the FP64 peak
        tf = fl["flops"] / (avg_ms / 1e3) / 1e12
        line["roofline"]["fp64"] = {"flops_per_launch": fl["flops"], "valu_flops": fl["valu_flops"],
                                    "mfma_flops": fl["mfma_flops"], "achieved": tf, "peak": FP64_PEAK_TFS,
                                    "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFS, "source": fl["source"]}
    tr = measured_traffic(args.workload, dom)
    if tr:
        line["roofline"]["traffic"] = tr["bytes"]
        line["roofline"]["traffic_detail"] = tr
    if not args.no_paths:
        line["paths"] = {"weno_fp64_riemann256": path_weno_riemann256(4 * args.path_steps),
                         "jet_fp32_bulk": path_jet(args.path_steps),
                         "ensemble256": path_ensemble(args.ensemble_steps),
                         "dg2_learned2048": path_learned(args.path_steps, args.learned_problems)}
    if cpu is not None:
        line["cpu_baseline"] = cpu
    if args.calibrate_copy_mib > 0:
        line["calibration_copy_bytes"] = device_copy(args.calibrate_copy_mib << 20)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
