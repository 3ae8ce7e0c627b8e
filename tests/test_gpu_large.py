"""GPU: a DG state above 4 GiB per array.

Below 4 GiB the fused DG kernel runs with 32-bit element offsets fixed at
compile time; above it, the form with 64-bit offsets, and every kernel that
turns a cell index into an offset of the row-interleaved state (limiter, ghost
ring, wave speed, the conversion at the C ABI) works past 2^32 bytes.  The
oracle cannot run such a grid in test time, so the check is a property: with
initial data that vary along y alone and periodic x sides, every column of a
wide grid must evolve exactly as the columns of a narrow grid run through the
same kernels.
"""
import numpy as np
import pytest

import oracle
from zuikakucfd_amd import meshio
from zuikakucfd_amd import zk

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def y_only_problem(ny, nx):
    """Two states across y = ny/2 with smooth variations along y, the same in
    every column; copy ghosts above and below, periodic sides."""
    i = np.arange(ny, dtype=np.float64)[:, None] * np.ones((1, nx))
    top = (i >= ny // 2).astype(np.float64)
    m = meshio.riemann_2d(16)
    m.rho = 1.0 + 0.5 * top + 0.1 * np.sin(2 * np.pi * i / ny)
    m.u = 0.3 * np.ones((ny, nx))
    m.v = 0.2 * np.cos(2 * np.pi * i / ny)
    m.p = 1.0 + 0.4 * top
    b = meshio.ring_bctype(ny, nx)
    b[3:ny - 3, :3] = -2
    b[3:ny - 3, nx - 3:] = -2
    m.bctype = b
    m.ny, m.nx = ny, nx
    m.da = 1.0 / 64
    return m


def test_dg3_state_above_4gib_matches_narrow_grid(monkeypatch):
    """70 x 122,886 cells at P = 3: 4.4 GB per state array.  State, limiter
    flags, conserved means and wave speed match a 70 x 70 grid bit for bit
    after 3 x 4 steps (both on the 64-bit-offset form of the fused kernel)."""
    ny, nx_big, nx_small, P = 70, 122886, 70, 3
    nm = (P + 1) ** 2
    assert 4 * nm * ny * nx_big * 8 > 2 ** 32
    say(f"dg P=3 {ny} x {nx_big} ({4 * nm * ny * nx_big * 8 / 1e9:.2f} GB per state) against {ny} x {nx_small}")
    monkeypatch.setenv("ZK_FUSED_GENERIC", "1")  # the narrow grid on the same kernel form
    runs = []
    for nx in (nx_big, nx_small):
        m = y_only_problem(ny, nx)
        s = zk.Solver(method=zk.METHOD_DG, P=P, limiter=zk.LIMITER_TVB, tvb_M=50.0)
        s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
        runs.append((s, m))
    monkeypatch.delenv("ZK_FUSED_GENERIC")
    big, small = runs[0][0], runs[1][0]
    lam = small.max_speed()
    assert big.max_speed() == lam
    dt = 0.3 * oracle.dg_cfl_limit(P) * runs[1][1].da / lam
    for rep in range(3):
        for s in (big, small):
            s.step(dt, 4)
        say(f"dg P=3 above 4 GiB: after {4 * (rep + 1)} steps")
        Qs = small.get_state()
        Qb = big.get_state()
        assert np.isfinite(Qs).all()
        for c in range(4):
            for mo in range(nm):
                ref = Qs[c, mo, :, 3:4]
                assert (Qs[c, mo, :, 3:nx_small - 3] == ref).all(), (c, mo)  # the narrow grid itself
                assert (Qb[c, mo, :, 3:nx_big - 3] == ref).all(), (rep, c, mo)
        del Qb
        Ub = big.get_conserved()
        for c in range(4):
            assert (Ub[c, :, 3:nx_big - 3] == Qs[c, 0, :, 3:4]).all(), c
        fb, fs = big.get_flags(), small.get_flags()
        assert (fb[:, 3:nx_big - 3] == fs[:, 3:4]).all()
        assert big.max_speed() == small.max_speed()
    assert fs[3:ny - 3, 3:nx_small - 3].any(), "the limiter never acted: the test would not cover it"
    for s, _ in runs:
        s.close()
