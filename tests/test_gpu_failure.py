"""Failure detection on the GPU path: a state that stops being finite is
reported (ZK_ERR_NUMERIC, naming the variable, the cell and the step count)
instead of being stepped on and written out, as the reference does.

The failures here are data: a NaN loaded through zk_set_state, or a flow that
opens a vacuum under the reference scheme (no positivity option), where the
oracle goes non-finite at the same step.  Every kernel runs to its normal end.
"""
import re
import subprocess

import numpy as np
import pytest

import oracle
from zuikakucfd_amd import meshio
from zuikakucfd_amd import zk

pytestmark = pytest.mark.gpu

ROOT = __import__("pathlib").Path(__file__).resolve().parents[1]


def say(msg):
    print(f"[case] {msg}", flush=True)


def diverging_mesh(n=46):
    """Periodic flow whose diverging columns open a vacuum within ~13 steps."""
    m = meshio.riemann_2d(64)
    x = (np.arange(n) - 3 + 0.5) / (n - 6)
    X, Y = np.meshgrid(x, x)
    m.rho = np.ones((n, n)); m.p = np.full((n, n), 0.1)
    u = np.where(X < 0.5, -3.0, 3.0)
    m.u = np.where((X < 0.25) | (X > 0.75), -u, u)
    m.v = 0.3 * np.sin(2 * np.pi * Y)
    m.bctype = meshio.ring_bctype(n, n, -2)
    m.nx = m.ny = n
    m.da = 1.0 / (n - 6)
    m.gamma = 1.4
    return m


def first_nonfinite_step_oracle(m, dt, limit=60):
    o = oracle.FV(m.da, m.gamma, oracle.prim_to_cons(m.gamma, m.rho, m.u, m.v, m.p), m.bctype)
    for k in range(1, limit + 1):
        o.step(dt)
        if not np.isfinite(o.U).all():
            return k
    return None


@pytest.mark.parametrize("method", ["weno", "dg"])
def test_nan_in_state_is_reported_and_blocks_stepping(method):
    """A NaN planted in one cell is reported by the next synchronising call
    with its variable and cell; stepping is refused until a finite state is
    set, after which it works again."""
    say(f"planted NaN, {method}")
    m = meshio.riemann_2d(48)
    if method == "weno":
        s = zk.Solver(method=zk.METHOD_WENO)
    else:
        s = zk.Solver(method=zk.METHOD_DG, P=2, limiter=zk.LIMITER_TVB)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    good = s.get_state()
    bad = good.copy()
    bad[2, 0, 20, 30] = np.nan
    with pytest.raises(zk.ZkNumericError, match=r"U2 .*cell \(i=20, j=30\).*after step 0"):
        s.set_state(bad)
    with pytest.raises(zk.ZkNumericError):
        s.step(1e-4, 2)
    with pytest.raises(zk.ZkNumericError):
        s.get_conserved()
    s.set_state(good)
    s.step(1e-4, 2)
    s.sync()
    assert np.isfinite(s.get_conserved()).all()
    s.close()


def test_vacuum_reported_at_the_oracle_step():
    """The reference scheme (no positivity option) on a flow that opens a
    vacuum: stepping one step per call, the first synchronising call that
    fails is the step at which the oracle's state stops being finite, and the
    message carries that step count."""
    say("vacuum, fixed dt, step of failure")
    m = diverging_mesh()
    o = oracle.FV(m.da, m.gamma, oracle.prim_to_cons(m.gamma, m.rho, m.u, m.v, m.p), m.bctype)
    dt = 0.4 * m.da / o.max_speed()
    k_oracle = first_nonfinite_step_oracle(m, dt)
    assert k_oracle is not None
    s = zk.Solver(method=zk.METHOD_WENO)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    k_gpu, msg = None, ""
    for k in range(1, k_oracle + 5):
        s.step(dt, 1)
        try:
            s.sync()
        except zk.ZkNumericError as e:
            k_gpu, msg = k, str(e)
            break
    print(f"[vacuum] oracle step {k_oracle}, GPU step {k_gpu}: {msg}", flush=True)
    assert k_gpu == k_oracle
    assert re.search(rf"after step {k_oracle}\b", msg)
    s.close()


def test_step_cfl_stops_on_nonfinite():
    """CFL stepping through the same vacuum stops with ZK_ERR_NUMERIC instead
    of running all its steps with a NaN time step."""
    say("vacuum, step_cfl")
    m = diverging_mesh()
    s = zk.Solver(method=zk.METHOD_WENO)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    with pytest.raises(zk.ZkNumericError, match="non-finite|not finite"):
        s.step_cfl(0.4, 1e30, 200)
    with pytest.raises(zk.ZkNumericError):
        s.step_cfl(0.4, 1e30, 5)
    s.close()


def test_dg_step_cfl_above_limit_refused():
    """DG CFL above 0.9 / (2 (2P + 1)) is refused (ZK_ERR_CFL), as the
    reference refuses to run (src_DG/solver.cpp:245-260)."""
    say("dg cfl above limit")
    m = meshio.riemann_2d(32)
    s = zk.Solver(method=zk.METHOD_DG, P=2, limiter=zk.LIMITER_TVB)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    with pytest.raises(zk.ZkCflError, match="stable limit"):
        s.step_cfl(0.2, 1e30, 3)
    t, n = s.step_cfl(0.09, 1e30, 3)
    assert n == 3 and t > 0
    s.close()


def test_ensemble_reports_nonfinite_instance():
    """An ensemble instance that goes non-finite stops and is named by
    zk_ensemble_sync; its neighbour in the ensemble is unaffected."""
    say("ensemble non-finite instance")
    mv = diverging_mesh()
    mo = meshio.random_riemann(301, 40)
    E = zk.Ensemble(positivity=False)
    for m in (mv, mo):
        E.add(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype, cfl=0.4)
    E.finalize()
    E.step(40)
    with pytest.raises(zk.ZkNumericError, match=r"first: 0 \(step"):
        E.sync()
    U, t = E.get(1)
    s = zk.Solver(method=zk.METHOD_WENO)
    s.load(mo.da, mo.gamma, mo.rho, mo.u, mo.v, mo.p, mo.bctype)
    ts, _ = s.step_cfl(0.4, 1e30, 40)
    Us = s.get_conserved()
    assert abs(t - ts) <= 1e-13 * ts
    np.testing.assert_allclose(U, Us, rtol=1e-10, atol=1e-12)
    s.close()
    E.close()


def test_cli_exits_nonzero_on_vacuum(tmp_path):
    """The drop-in executable stops with a non-zero status and names the
    failure instead of writing non-finite result files."""
    say("cli vacuum")
    m = diverging_mesh()
    meshio.write_mesh_folder(m, tmp_path / "mesh")
    o = oracle.FV(m.da, m.gamma, oracle.prim_to_cons(m.gamma, m.rho, m.u, m.v, m.p), m.bctype)
    dt = 0.4 * m.da / o.max_speed()
    exe = ROOT / "zuikakucfd_amd" / "bin" / "zk_weno"
    r = subprocess.run([str(exe), str(tmp_path / "mesh"), repr(dt), "60", "--quiet"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode != 0
    assert "non-finite" in r.stderr
