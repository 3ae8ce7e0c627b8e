"""Autotuner for kernel variants (tile shapes, occupancy variants).

For a given problem the tuner builds a throwaway solver (or ensemble) from a
factory, times every candidate variant over a few steps with HIP events on
the library's own stream, and returns the fastest.  Results are cached in a
JSON file keyed by method, degree, precision and problem size
($ZK_AUTOTUNE_CACHE, default ~/.cache/zuikakucfd_amd/autotune.json); a cache
that cannot be written is skipped silently.
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Callable

from . import zk


def cache_path() -> Path:
    p = os.environ.get("ZK_AUTOTUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "zuikakucfd_amd" / "autotune.json"


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except Exception:
        return {}


def _store(key: str, value) -> None:
    try:
        d = _load()
        d[key] = value
        cache_path().parent.mkdir(parents=True, exist_ok=True)
        cache_path().write_text(json.dumps(d, indent=1, sort_keys=True))
    except Exception:
        pass


def solver_candidates(method: int, P: int) -> list[tuple[int, int]]:
    """(fv_tile, dg_variant) pairs worth timing for a single-grid solver."""
    if method == zk.METHOD_WENO:
        return [(1, 0), (2, 0)]
    if P == 3:
        return [(0, 0), (0, 117), (0, 134), (0, 26), (0, 2)]
    return [(0, 0), (0, 117), (0, -1)]


def tune_solver(factory: Callable[[], zk.Solver], dt: float, key: str, steps: int = 4,
                use_cache: bool = True) -> tuple[tuple[int, int], dict]:
    """Time every candidate on a fresh solver from `factory`; return the best
    (fv_tile, dg_variant) and the timings in ms per step."""
    if use_cache:
        hit = _load().get(key)
        if hit:
            return tuple(hit["best"]), hit["ms_per_step"]
    s0 = factory()
    cands = solver_candidates(s0.method, s0.P)
    s0.close()
    timings = {}
    for fv_tile, dgv in cands:
        s = factory()
        s.set_tuning(fv_tile, dgv)
        s.step(dt, 1)  # first launch compiles nothing but warms caches
        s.sync()
        ms = s.step_timed(dt, steps)
        timings[f"{fv_tile},{dgv}"] = ms / steps
        s.close()
    best = min(timings, key=timings.get)
    choice = tuple(int(x) for x in best.split(","))
    _store(key, {"best": list(choice), "ms_per_step": timings})
    return choice, timings


def tune_ensemble(factory: Callable[[], zk.Ensemble], key: str, steps: int = 3,
                  use_cache: bool = True) -> tuple[int, dict]:
    """Time the ensemble work-item tile shapes (1 = 16x15, 2 = 32x15) with
    graph replay; return the fastest and the timings in ms per step."""
    if use_cache:
        hit = _load().get(key)
        if hit:
            return int(hit["best"]), hit["ms_per_step"]
    timings = {}
    for tile in (1, 2):
        E = factory()
        E.set_tuning(tile=tile, use_graph=True)
        E.step(1)
        E.sync()
        t0 = time.perf_counter()
        E.step(steps)
        E.sync()
        timings[str(tile)] = (time.perf_counter() - t0) * 1e3 / steps
        E.close()
    best = int(min(timings, key=timings.get))
    _store(key, {"best": best, "ms_per_step": timings})
    return best, timings
