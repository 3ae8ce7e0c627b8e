"""Learned troubled-cell indicator for DG P=2: harvest, train, evaluate.

The network (zk.Mlp) is trained on (normalised neighbourhood modal
coefficients, TVB flag) pairs that the GPU solver itself produces while it
runs the plain TVB limiter on seeded random four-quadrant Riemann problems.
In use (ZK_LIMITER_LEARNED) only candidate cells are scored: those with a
linear mode above M h^2, since every other cell passes the TVB test
unchanged.  Scores above `hi` mark a troubled cell, below `lo` a smooth one;
scores in between and a random audit fraction fall back to the TVB test,
which also counts the network's misses and false flags.
"""
from __future__ import annotations

import time
from pathlib import Path

import numpy as np

from . import meshio, zk

DATA = Path(__file__).resolve().parent / "data"
WEIGHTS = DATA / "troubled_cell_p2.npz"
CFL_P2 = 0.9 / (2.0 * (2 * 2 + 1))


def dg_solver(m, limiter: int, M: float = 50.0, positivity: bool = False) -> zk.Solver:
    s = zk.Solver(method=zk.METHOD_DG, P=2, limiter=limiter, tvb_M=M, positivity=positivity)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    return s


def harvest(mlp: zk.Mlp, seeds, n: int = 96, steps: int = 60, keep_pos: float = 1.0,
            keep_neg: float = 1.0) -> int:
    """Run plain TVB DG P=2 on each seeded problem, appending the (features,
    flag) pairs of candidate cells to mlp's set (positives kept with
    probability keep_pos, negatives keep_neg)."""
    for seed in seeds:
        m = meshio.random_riemann(int(seed), n)
        s = dg_solver(m, zk.LIMITER_TVB)
        dt = CFL_P2 * m.da / max(s.max_speed(), 1e-7)
        s.harvest(mlp, keep_pos, keep_neg)
        s.step(dt, steps)
        s.sync()
        s.harvest(None)
        s.close()
    return mlp.dataset(0)[0]


def train(mlp: zk.Mlp, epochs: int = 30, batch: int = 256, lr: float = 2e-3, seed: int = 0) -> np.ndarray:
    return mlp.train(epochs, batch, lr, seed)


def save(mlp: zk.Mlp, path=WEIGHTS, **meta) -> None:
    """Parameters and training metadata as .npz, plus the raw little-endian
    float32 vector (.f32) that the zk_dg3 executable's --weights reads."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    w = mlp.get_params()
    np.savez(path, w=w, **{k: np.asarray(v) for k, v in meta.items()})
    w.astype("<f4").tofile(path.with_suffix(".f32"))


def load(path=WEIGHTS, device: int = 0) -> zk.Mlp:
    d = np.load(path)
    mlp = zk.Mlp(device=device)
    mlp.set_params(d["w"])
    return mlp


def timed_run(s: zk.Solver, dt: float, steps: int, chunk: int = 500) -> tuple[float, int | None]:
    """Advance `steps` steps in timed chunks (HIP events on the handle's
    stream); (milliseconds, None), or (milliseconds so far, the step count of
    the check that found a non-finite state)."""
    ms, done = 0.0, 0
    while done < steps:
        k = min(chunk, steps - done)
        try:
            ms += s.step_timed(dt, k)
        except zk.ZkNumericError:
            return ms, done + k
        done += k
    return ms, None


def evaluate(mlp: zk.Mlp, seed: int, n: int, steps: int, lo: float = 0.2, hi: float = 0.8,
             audit: float = 0.02, positivity: bool = False) -> dict:
    """Plain TVB and learned runs of one problem from the same state with the
    same fixed dt: rates, density L1 difference, the network's miss and
    false-flag rates on audited cells and the share of cells in the band.
    positivity: both runs with the DG positivity option."""
    m = meshio.random_riemann(int(seed), n)
    plain = dg_solver(m, zk.LIMITER_TVB, positivity=positivity)
    dt = CFL_P2 * m.da / max(plain.max_speed(), 1e-7)
    learned = dg_solver(m, zk.LIMITER_LEARNED, positivity=positivity)
    learned.attach_mlp(mlp, lo, hi, audit)
    ncell = plain.updated_cells()
    out = {"seed": int(seed), "n": n, "steps": steps, "dt": dt, "lo": lo, "hi": hi, "audit": audit,
           "positivity": positivity}
    for name, s in (("tvb", plain), ("learned", learned)):
        s.step(dt, 2)
        ms, fail = timed_run(s, dt, steps)
        out[f"{name}_cell_stage_per_s"] = 3 * ncell * steps / (ms / 1e3) if fail is None else None
        out[f"{name}_nonfinite_step"] = fail
    if out["tvb_nonfinite_step"] or out["learned_nonfinite_step"]:
        out["finite"] = False
        plain.close()
        learned.close()
        return out
    Up, Ul = plain.get_conserved(), learned.get_conserved()
    mask = np.zeros_like(m.bctype, dtype=bool)
    mask[3:-3, 3:-3] = True
    out["l1_rho_vs_tvb"] = float(np.abs(Ul[0][mask] - Up[0][mask]).sum() / np.abs(Up[0][mask]).sum())
    st = learned.learned_stats()
    out.update(st)
    out.update(rates(st))
    out["candidate_fraction"] = st["candidates"] / max(1, 3 * (steps + 2) * ncell)
    out["band_fraction"] = st["in_band"] / max(1, st["candidates"])
    out["finite"] = bool(np.isfinite(Ul).all())
    plain.close()
    learned.close()
    return out


def rates(st: dict) -> dict:
    """Classification rates of the network's 0.5 decision against the TVB
    test over the audited candidate cells: agreement, recall of troubled
    cells, specificity on smooth ones and their mean (balanced accuracy,
    which a constant answer scores 0.5 on, whatever the class balance)."""
    a = max(st["audited"], 1)
    pos = st["audited_positive"]
    neg = st["audited"] - pos
    recall = (pos - st["miss"]) / pos if pos else 1.0
    spec = (neg - st["false_flag"]) / neg if neg else 1.0
    return {"agree_rate": st["agree"] / a, "miss_rate": st["miss"] / a, "false_flag_rate": st["false_flag"] / a,
            "recall": recall, "specificity": spec, "balanced_accuracy": 0.5 * (recall + spec),
            "positive_share": pos / a}


def agreement(mlp: zk.Mlp, seed: int, n: int, steps: int) -> dict:
    """The network's decisions against the TVB test on every candidate cell
    of a run (audit fraction 1: every limiting decision is the TVB test's)."""
    m = meshio.random_riemann(int(seed), n)
    s = dg_solver(m, zk.LIMITER_LEARNED)
    s.attach_mlp(mlp, 0.0, 1.0, 1.0)
    dt = CFL_P2 * m.da / max(s.max_speed(), 1e-7)
    s.step(dt, steps)
    st = s.learned_stats()
    s.close()
    return {"audited": st["audited"], **rates(st)}


def c4_problem(mlp: zk.Mlp, seed: int, n: int = 2048, steps: int = 2000) -> dict:
    """One seeded four-quadrant DG P=2 problem run three ways for `steps`
    fixed-dt steps at nu = CFL_P2: the reference scheme (TVB alone), then TVB
    and the learned indicator, both with the DG positivity option (evaluate).
    Non-finite states are found by the failure check every 500 steps."""
    m = meshio.random_riemann(int(seed), n)
    s = dg_solver(m, zk.LIMITER_TVB)
    dt = CFL_P2 * m.da / max(s.max_speed(), 1e-7)
    s.step(dt, 2)
    ms, fail = timed_run(s, dt, steps)
    ncell = s.updated_cells()
    s.close()
    out = evaluate(mlp, seed, n, steps, positivity=True)
    out["plain_tvb_nonfinite_step"] = fail
    out["plain_tvb_cell_stage_per_s"] = 3 * ncell * steps / (ms / 1e3) if fail is None else None
    return out


def build_default(seeds=range(0, 48), n: int = 96, steps: int = 60, epochs: int = 40, capacity: int = 400000,
                  path=WEIGHTS) -> dict:
    """Harvest, train and save the network shipped with the package."""
    mlp = zk.Mlp(seed=1234, capacity=capacity)
    count = harvest(mlp, seeds, n=n, steps=steps)
    losses = train(mlp, epochs=epochs)
    _, _, y = mlp.dataset(count)
    save(mlp, path, losses=losses, samples=count, positives=int(y.sum()), seeds=list(seeds), n=n, steps=steps)
    return {"samples": count, "positives": int(y.sum()), "loss_first": float(losses[0]), "loss_last": float(losses[-1])}


def main(argv=None) -> int:
    import argparse
    import json

    ap = argparse.ArgumentParser(description="harvest, train and save the DG P=2 troubled-cell network")
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--n", type=int, default=2048, help="grid size of the harvest problems")
    ap.add_argument("--seeds", type=int, nargs=2, default=(100, 112), help="harvest seeds [a, b)")
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--keep-pos", type=float, default=0.03, help="share of troubled candidate cells kept")
    ap.add_argument("--keep-neg", type=float, default=0.5, help="share of smooth candidate cells kept")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--capacity", type=int, default=1_500_000)
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[0, 1])
    ap.add_argument("--eval-steps", type=int, default=20)
    ap.add_argument("--c4", type=int, nargs=2, metavar=("A", "B"),
                    help="instead of training: run c4_problem on seeds [A, B) with the shipped network, "
                         "one JSON line per seed (--n, --c4-steps)")
    ap.add_argument("--c4-steps", type=int, default=2000)
    a = ap.parse_args(argv)
    if a.c4:
        mlp = load()
        for seed in range(*a.c4):
            print(json.dumps(c4_problem(mlp, seed, a.n, a.c4_steps)), flush=True)
        mlp.close()
        return 0
    mlp = zk.Mlp(seed=1234, capacity=a.capacity)
    t0 = time.perf_counter()
    for seed in range(*a.seeds):
        count = harvest(mlp, [seed], n=a.n, steps=a.steps, keep_pos=a.keep_pos, keep_neg=a.keep_neg)
        print(f"[harvest] seed {seed}: {count} samples ({time.perf_counter() - t0:.1f} s)", flush=True)
    _, _, y = mlp.dataset(count)
    losses = train(mlp, epochs=a.epochs)
    print(f"[train] {count} samples, {int(y.sum())} positive; loss {losses[0]:.4f} -> {losses[-1]:.4f} "
          f"({time.perf_counter() - t0:.1f} s)", flush=True)
    save(mlp, a.out, losses=losses, samples=count, positives=int(y.sum()), seeds=list(range(*a.seeds)), n=a.n,
         steps=a.steps, keep_pos=a.keep_pos, keep_neg=a.keep_neg)
    for seed in a.eval_seeds:
        print(json.dumps({"agreement": agreement(mlp, seed, a.n, a.eval_steps)}), flush=True)
        print(json.dumps({"evaluate": evaluate(mlp, seed, a.n, a.eval_steps)}), flush=True)
    mlp.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
