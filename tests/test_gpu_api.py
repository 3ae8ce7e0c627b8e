"""The C ABI's contract on the GPU: documented errors, degenerate grids
(one updated cell, none) against the oracle, and size-independent
properties at the headline sizes (exact conservation on periodic data)."""
import numpy as np
import pytest

import oracle
from helpers import rel_l1
from zuikakucfd_amd import meshio, zk

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def tiny(n=7, seed=0, center=0):
    rng = np.random.default_rng(seed)
    m = meshio.riemann_2d(64)
    m.rho = 0.5 + rng.random((n, n)); m.u = rng.uniform(-1, 1, (n, n))
    m.v = rng.uniform(-1, 1, (n, n)); m.p = 0.5 + rng.random((n, n))
    m.bctype = meshio.ring_bctype(n, n)
    m.bctype[3, 3] = center
    m.nx = m.ny = n
    m.da = 0.1
    return m


def test_documented_errors():
    say("api errors")
    with pytest.raises(zk.ZkError, match="degree"):
        zk.Solver(method=zk.METHOD_DG, P=4)
    with pytest.raises(zk.ZkError, match="P=2"):
        zk.Solver(method=zk.METHOD_DG, P=3, limiter=zk.LIMITER_LEARNED)
    s = zk.Solver(method=zk.METHOD_WENO)
    with pytest.raises(zk.ZkError, match="not loaded"):
        s.step(1e-4, 1)
    m = tiny(6)
    with pytest.raises(zk.ZkError, match=">= 7"):
        s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    m = tiny()
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    with pytest.raises(zk.ZkError, match="dt >= 0"):
        s.step(-1.0, 1)
    with pytest.raises(zk.ZkError, match="fv_tile"):
        s.set_tuning(3, 0)
    with pytest.raises(zk.ZkError, match="dg_variant"):
        s.set_tuning(0, 5)
    with pytest.raises(zk.ZkError, match="ZK_RECON"):
        s.set_reconstruction(4)
    s.close()
    d = zk.Solver(method=zk.METHOD_DG, P=2, limiter=zk.LIMITER_LEARNED)
    d.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    with pytest.raises(zk.ZkError, match="network"):
        d.step(1e-4, 1)
    with pytest.raises(zk.ZkError, match="WENO handle"):
        d.set_reconstruction(zk.RECON_MUSCL_MINMOD)
    d.close()
    t = zk.Solver(method=zk.METHOD_DG, P=2, limiter=zk.LIMITER_TVB)
    t.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    net = zk.Mlp(seed=1)
    with pytest.raises(zk.ZkError, match="LEARNED"):
        t.attach_mlp(net, 0.2, 0.8, 0.0)
    t.close()


def test_stream_bandwidth():
    """zk_stream_bandwidth: positive rates below the HBM peak for 256 MiB
    arrays, and its argument errors."""
    say("stream bandwidth reference")
    copy, add = zk.stream_bandwidth(256 << 20)
    assert 500.0 < copy < 8000.0 and 500.0 < add < 8000.0, (copy, add)
    with pytest.raises(zk.ZkError, match="4 KiB"):
        zk.stream_bandwidth(1024)
    with pytest.raises(zk.ZkError, match="device ordinal"):
        zk.stream_bandwidth(1 << 20, device=99)


@pytest.mark.parametrize("center", [0, -1])
def test_weno_degenerate_grids(center):
    """7x7: one updated cell, or none (the centre a ghost): only the stage
    combination of the ring, exactly as the oracle."""
    say(f"weno 7x7 centre={center}")
    m = tiny(center=center)
    s = zk.Solver(method=zk.METHOD_WENO)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    assert s.updated_cells() == (1 if center == 0 else 0)
    o = oracle.FV(m.da, m.gamma, oracle.prim_to_cons(m.gamma, m.rho, m.u, m.v, m.p), m.bctype)
    s.step(1e-3, 4)
    o.step(1e-3, 4)
    U = s.get_conserved()
    for c in range(4):
        assert rel_l1(U[c], o.U[c]) <= 1e-12, c


@pytest.mark.parametrize("P", [1, 2, 3])
@pytest.mark.parametrize("variant", [0, 104, -1])
@pytest.mark.parametrize("center", [0, -1])
def test_dg_degenerate_grids(P, variant, center):
    say(f"dg P={P} 7x7 variant={variant} centre={center}")
    m = tiny(seed=P, center=center)
    s = zk.Solver(method=zk.METHOD_DG, P=P, limiter=zk.LIMITER_TVB)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    s.set_tuning(0, variant)
    o = oracle.DG.from_prims(P, m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype, limiter=2, tvb_M=50.0)
    s.step(1e-3, 3)
    o.step(1e-3, 3)
    Q = s.get_state()
    for c in range(4):
        assert rel_l1(Q[c], o.Q[c]) <= 1e-12, c
    np.testing.assert_array_equal(s.get_flags(), o.flags)


def test_dg3_conservation_at_1024():
    """Headline size: DG P=3 with TVB on the periodic smooth wave (L2-projected
    modes): mass, momentum and energy are conserved to rounding."""
    say("dg P=3 1024^2 periodic conservation")
    n = 1024
    m = meshio.smooth_wave(n)
    s = zk.Solver(method=zk.METHOD_DG, P=3, limiter=zk.LIMITER_TVB)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    s.set_state(meshio.smooth_wave_modes(n, 3))
    Q0 = s.get_state()[:, 0, 3:-3, 3:-3].sum(axis=(1, 2))
    dt = 0.9 / 14 * m.da / s.max_speed()
    s.step(dt, 5)
    Q = s.get_state()
    assert np.isfinite(Q).all()
    Q1 = Q[:, 0, 3:-3, 3:-3].sum(axis=(1, 2))
    np.testing.assert_allclose(Q1, Q0, rtol=1e-12, atol=1e-12 * np.abs(Q0).max())
    assert s.get_flags().sum() == 0  # smooth: nothing limited


@pytest.mark.parametrize("prec", [zk.PREC_FP64, zk.PREC_FP32_BULK])
def test_weno_conservation_at_2048(prec):
    say(f"weno 2048^2 periodic conservation prec={prec}")
    n = 2048
    m = meshio.smooth_wave(n)
    s = zk.Solver(method=zk.METHOD_WENO, precision=prec)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    U0 = s.get_conserved()[:, 3:-3, 3:-3].sum(axis=(1, 2))
    dt = 0.5 * m.da / s.max_speed()
    s.step(dt, 10)
    U = s.get_conserved()
    assert np.isfinite(U).all()
    U1 = U[:, 3:-3, 3:-3].sum(axis=(1, 2))
    np.testing.assert_allclose(U1, U0, rtol=1e-12, atol=1e-12 * np.abs(U0).max())


@pytest.mark.parametrize("case", ["dg3_tvb_positivity", "dg2_tvb", "weno_jet_fp32", "ensemble_fp64"])
def test_runs_are_bitwise_deterministic(case):
    """Two fresh handles on the same input give the same state bit for bit.
    The only atomics on these paths are integer ones (the wave-speed maximum on
    the ordered bits of positive doubles, the TVB pass's list of uncertified
    cells, whose order varies but whose cells are each scaled by one wave,
    and the failure record's minima)."""
    say(f"determinism {case}")

    def run():
        if case == "ensemble_fp64":
            E = zk.Ensemble(precision=zk.PREC_FP64, positivity=True)
            for k, n in enumerate([40, 56, 72]):
                m = meshio.random_riemann(300 + k, n)
                E.add(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype, cfl=0.5)
            E.finalize()
            E.step(30)
            E.sync()
            out = np.concatenate([E.get(k)[0].ravel() for k in range(3)])
            E.close()
            return out
        if case == "weno_jet_fp32":
            m = meshio.jet(512, 256)
            s = zk.Solver(method=zk.METHOD_WENO, precision=zk.PREC_FP32_BULK, positivity=True)
            s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
            dt = 0.25 * m.da / s.max_speed()
            s.step(dt, 40)
        else:
            P = 3 if case.startswith("dg3") else 2
            # seed 6 forms a near-vacuum region (the reference scheme fails there
            # by step 120), where cells fall outside the positivity certificate
            m = meshio.random_riemann(6 if case.endswith("positivity") else 1, 96)
            s = zk.Solver(method=zk.METHOD_DG, P=P, limiter=zk.LIMITER_TVB, positivity=case.endswith("positivity"))
            s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
            dt = 0.5 * oracle.dg_cfl_limit(P) * m.da / s.max_speed()
            s.step(dt, 60)
        out = s.get_state().ravel()
        s.close()
        return out

    a, b = run(), run()
    assert np.isfinite(a).all()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
