"""Pins of the MLP oracle: float64 gradients against central differences,
and the BF16 rounding helper against closed cases."""
import numpy as np
import pytest

import mlp_ref


def random_params(rng, scale=0.3):
    w = rng.normal(size=mlp_ref.P_N) * scale
    w[mlp_ref.P_W1:mlp_ref.P_B1].reshape(64, 192)[:, 180:] = 0.0
    return w


def test_bf16_rounding_closed_cases():
    assert mlp_ref.bf16_round(1.0) == 1.0
    # 1 + 2^-8 is exactly halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert mlp_ref.bf16_round(1.0 + 2.0 ** -8) == 1.0
    assert mlp_ref.bf16_round(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert mlp_ref.bf16_round(-2.5) == -2.5
    b = mlp_ref.bf16_bits(np.array([1.0, -2.0, 0.5], dtype=np.float32))
    np.testing.assert_array_equal(b, [0x3F80, 0xC000, 0x3F00])
    np.testing.assert_array_equal(mlp_ref.bits_to_float(b), [1.0, -2.0, 0.5])


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    w = random_params(rng)
    X = rng.normal(size=(8, 192))
    X[:, 180:] = 0.0
    y = (rng.random(8) < 0.5).astype(float)
    loss, g, _ = mlp_ref.forward_backward(w, X, y, rnd=False)
    idx = np.concatenate([rng.integers(mlp_ref.P_W1, mlp_ref.P_W1 + 64 * 180, 12),
                          rng.integers(mlp_ref.P_B1, mlp_ref.P_W2, 4), rng.integers(mlp_ref.P_W2, mlp_ref.P_B2, 8),
                          rng.integers(mlp_ref.P_B2, mlp_ref.P_W3, 4), rng.integers(mlp_ref.P_W3, mlp_ref.P_B3, 4),
                          [mlp_ref.P_B3]])
    for k in idx:
        e = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[k] += e
        wm[k] -= e
        fd = (mlp_ref.forward_backward(wp, X, y, rnd=False)[0] - mlp_ref.forward_backward(wm, X, y, rnd=False)[0]) / (2 * e)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_loss_is_bce_of_sigmoid():
    rng = np.random.default_rng(1)
    w = np.zeros(mlp_ref.P_N)
    w[mlp_ref.P_B3] = 0.7  # constant logit
    X = rng.normal(size=(5, 192))
    y = np.array([1, 0, 1, 1, 0], dtype=float)
    loss, g, p = mlp_ref.forward_backward(w, X, y, rnd=False)
    s = 1 / (1 + np.exp(-0.7))
    assert loss == pytest.approx(np.mean(-(y * np.log(s) + (1 - y) * np.log(1 - s))), rel=1e-12)
    np.testing.assert_allclose(p, s)
    assert g[mlp_ref.P_B3] == pytest.approx(np.mean(s - y), rel=1e-12)


def test_shipped_network_files():
    """The shipped DG P=2 network: the .npz (with its training metadata) and
    the raw float32 vector the zk_dg3 executable reads hold the same
    parameters, and the oracle's scores with them are probabilities."""
    from pathlib import Path

    data = Path(__file__).resolve().parents[1] / "zuikakucfd_amd" / "data"
    d = np.load(data / "troubled_cell_p2.npz")
    w = d["w"]
    assert w.shape == (mlp_ref.P_N,) and w.dtype == np.float32
    raw = np.fromfile(data / "troubled_cell_p2.f32", dtype="<f4")
    np.testing.assert_array_equal(raw, w)
    assert int(d["samples"]) > 0 and 0 < int(d["positives"]) < int(d["samples"])
    assert d["losses"][-1] < d["losses"][0]
    # the pad columns of W1 (features 180..191) were never trained away from 0
    W1 = w[mlp_ref.P_W1:mlp_ref.P_B1].reshape(mlp_ref.H, mlp_ref.IN)
    assert np.all(W1[:, mlp_ref.FEAT:] == 0)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(32, mlp_ref.IN)) * 0.3
    X[:, mlp_ref.FEAT:] = 0
    s = mlp_ref.scores(w.astype(np.float64), X)
    assert np.isfinite(s).all() and (s >= 0).all() and (s <= 1).all()
