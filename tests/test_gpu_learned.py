"""The learned troubled-cell indicator on the GPU: gradients and scores
against the CPU oracle, features against the oracle's restatement, the audit
path against the plain TVB limiter, and what training gains."""
import numpy as np
import pytest

import mlp_ref
import oracle
from helpers import rel_l1
from zuikakucfd_amd import learned, meshio, zk

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def random_features(rng, n):
    X = rng.normal(size=(n, 192)) * 0.5
    X[:, 180:] = 0.0
    return mlp_ref.bf16_bits(X)


def test_gradients_match_oracle():
    say("mlp gradients vs oracle")
    rng = np.random.default_rng(11)
    net = zk.Mlp(seed=5)
    w = net.get_params().astype(np.float64)
    Xb = random_features(rng, 96)
    y = (rng.random(96) < 0.4).astype(np.float32)
    g, loss = net.grad(Xb, y)
    lref, gref, _ = mlp_ref.forward_backward(w, mlp_ref.bits_to_float(Xb), y.astype(np.float64))
    assert loss == pytest.approx(lref, rel=1e-4)
    blocks = {"W1": (mlp_ref.P_W1, mlp_ref.P_B1), "b1": (mlp_ref.P_B1, mlp_ref.P_W2),
              "W2": (mlp_ref.P_W2, mlp_ref.P_B2), "b2": (mlp_ref.P_B2, mlp_ref.P_W3),
              "w3": (mlp_ref.P_W3, mlp_ref.P_B3), "b3": (mlp_ref.P_B3, mlp_ref.P_N)}
    for name, (a, b) in blocks.items():
        err = np.linalg.norm(g[a:b] - gref[a:b]) / max(np.linalg.norm(gref[a:b]), 1e-30)
        assert err < 2e-3, (name, err)


def test_scores_match_oracle():
    say("mlp inference vs oracle")
    rng = np.random.default_rng(12)
    net = zk.Mlp(seed=6)
    w = net.get_params().astype(np.float64)
    Xb = random_features(rng, 64)
    s = net.infer(Xb)
    np.testing.assert_allclose(s, mlp_ref.scores(w, mlp_ref.bits_to_float(Xb)), atol=2e-5)


def test_features_and_fused_scores():
    """Features from the device state equal the oracle's restatement bit for
    bit; the fused score path equals inference on those features."""
    say("mlp features from a DG P=2 state")
    m = meshio.random_riemann(5, 24)
    s = learned.dg_solver(m, zk.LIMITER_TVB)
    s.step(2e-4, 3)
    Q = s.get_state()
    o = oracle.DG(2, m.da, m.gamma, Q, m.bctype)
    o.fill()
    ij = [(i, j) for i in range(3, 21) for j in range(3, 21)][:96]
    cells = np.array([i * m.nx + j for i, j in ij], dtype=np.int32)
    net = zk.Mlp(seed=7)
    X, sc = s.mlp_features(cells, net)
    Xref = mlp_ref.bf16_bits(mlp_ref.features(o.Q, ij))
    np.testing.assert_array_equal(X, Xref)
    np.testing.assert_allclose(sc, net.infer(X), rtol=0, atol=1e-6)


def test_full_audit_equals_tvb():
    """With every cell audited the learned limiter takes every decision from
    the TVB test: state and flags equal the plain TVB run."""
    say("learned limiter, audit 1.0, vs TVB")
    m = meshio.random_riemann(21, 40)
    a = learned.dg_solver(m, zk.LIMITER_TVB)
    b = learned.dg_solver(m, zk.LIMITER_LEARNED)
    net = zk.Mlp(seed=8)
    b.attach_mlp(net, 0.3, 0.7, 1.0)
    dt = learned.CFL_P2 * m.da / a.max_speed()
    a.step(dt, 8)
    b.step(dt, 8)
    Qa, Qb = a.get_state(), b.get_state()
    for c in range(4):
        assert rel_l1(Qb[c], Qa[c]) <= 1e-13
    np.testing.assert_array_equal(a.get_flags(), b.get_flags())
    st = b.learned_stats()
    assert st["audited"] == st["agree"] + st["miss"] + st["false_flag"] > 0
    assert st["audited"] == st["candidates"]  # audit 1.0: every candidate is audited


def test_training_gains_over_untrained():
    """Harvest candidate cells from the solver's own TVB runs and train on the
    GPU.  On a held-out problem the trained network separates troubled from
    smooth candidates far better than the untrained one (balanced accuracy,
    on which any constant answer scores 0.5), and its learned-limiter run
    stays close to plain TVB."""
    say("harvest + train + evaluate")
    net = zk.Mlp(seed=9, capacity=200000)
    count = learned.harvest(net, seeds=range(300, 306), n=96, steps=30, keep_pos=0.25, keep_neg=1.0)
    assert count > 5000
    _, _, y = net.dataset(count)
    assert 0 < y.sum() < count
    base = zk.Mlp(seed=9)
    before = learned.agreement(base, seed=999, n=96, steps=20)
    losses = net.train(epochs=40, batch=128, lr=2e-3)
    assert losses[-1] < 0.7 * losses[0]
    after = learned.agreement(net, seed=999, n=96, steps=20)
    print(f"[learned] before {before} after {after}", flush=True)
    assert after["balanced_accuracy"] > before["balanced_accuracy"] + 0.1
    assert after["recall"] > 0.8
    ev = learned.evaluate(net, seed=999, n=96, steps=20)
    assert ev["finite"] and ev["l1_rho_vs_tvb"] < 1e-2


def test_shipped_network_against_plain_tvb():
    """What the shipped network gains over the method without it, on held-out
    seeds of the C4 draws at 512^2 (60 steps), against plain TVB from the same
    state and dt.  It gains nothing measurable: its decisions on audited
    candidates beat a constant answer (balanced accuracy above 0.5, recall
    above 0.7), but its density departs from plain TVB by about 4e-4 (relative
    L1; bounded here at 1e-3, above the 1e-4 the chosen forms aim at) and its
    run is slower, since the TVB test reads four neighbour means per candidate
    and the network spends 31 kflop.  The rates are printed, not asserted."""
    say("shipped network vs plain TVB, seeds 60..62 at 512^2")
    net = learned.load()
    rows = [learned.evaluate(net, seed=s, n=512, steps=60) for s in (60, 61, 62)]
    for r in rows:
        print(f"[learned vs tvb] seed {r['seed']}: L1(rho) {r['l1_rho_vs_tvb']:.2e}, "
              f"rate {r['learned_cell_stage_per_s'] / r['tvb_cell_stage_per_s']:.2f}x of TVB, "
              f"balanced accuracy {r['balanced_accuracy']:.2f}, recall {r['recall']:.2f}", flush=True)
        assert r["finite"]
        assert r["l1_rho_vs_tvb"] <= 1e-3
        assert r["balanced_accuracy"] > 0.5
        assert r["recall"] > 0.7
    net.close()
