"""CPU oracle of the troubled-cell MLP (test infrastructure only).

Restates, in numpy, the network of the learned indicator and the features it
reads, with the same rounding points as the GPU kernels: features, W1, W2,
the first hidden layer and the two back-propagated hidden gradients are
rounded to BF16 (round to nearest even) before they enter a product; all
products and sums are accumulated in float64 here (FP32 on the GPU), so
results agree to FP32-accumulation error.  With rounding switched off the same
code is an exact float64 network, which the finite-difference test pins.

Parameter layout (FP32 vector): W1[64][192], b1[64], W2[64][64], b2[64],
w3[64], b3.  Features of a cell: k = n*36 + c*9 + m for neighbour n in (self,
E = j+1, W = j-1, S = i+1, N = i-1), variable c, mode m, taken from the
ghost-filled stage state, divided by s_c = |mean_c| + |mean_rho| + 1e-12 of
the cell, with the cell's own means subtracted from mode 0; 12 zero pads.
"""
from __future__ import annotations

import numpy as np

IN, FEAT, H = 192, 180, 64
P_W1 = 0
P_B1 = P_W1 + H * IN
P_W2 = P_B1 + H
P_B2 = P_W2 + H * H
P_W3 = P_B2 + H
P_B3 = P_W3 + H
P_N = P_B3 + 1


def bf16_round(x):
    """float -> nearest-even BF16 value (as float64); float32 is rounded first,
    as a device conversion from double goes through float."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    b = ((u + r) >> 16) & 0xFFFF
    return (b << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x) -> np.ndarray:
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    return (((u + r) >> 16) & 0xFFFF).astype(np.uint16)


def bits_to_float(b) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[P_W1:P_B1].reshape(H, IN), w[P_B1:P_W2], w[P_W2:P_B2].reshape(H, H), w[P_B2:P_W3], w[P_W3:P_B3],
            w[P_B3])


def forward_backward(w, X, y, rnd: bool = True):
    """Loss (mean BCE of sigmoid outputs) and gradient vector for samples
    X[n][192] (feature values) with labels y[n]."""
    R = bf16_round if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(w)
    W1b, W2b = R(W1), R(W2)
    X = R(X)
    n = X.shape[0]
    h1pre = W1b @ X.T                       # [64][n]
    a1 = h1pre + b1[:, None]
    h1 = R(np.maximum(a1, 0.0))
    h2pre = W2b @ h1
    a2 = h2pre + b2[:, None]
    h2 = np.maximum(a2, 0.0)
    z = w3 @ h2 + b3
    ez = np.exp(-np.abs(z))  # sigmoid without overflow for either sign of z
    p = np.where(z >= 0, 1.0 / (1.0 + ez), ez / (1.0 + ez))
    loss = float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    dz = (p - y) / n
    g = np.zeros(P_N)
    g[P_W3:P_B3] = h2 @ dz
    g[P_B3] = dz.sum()
    dh2_raw = np.where(a2 > 0, dz[None, :] * w3[:, None], 0.0)
    g[P_B2:P_W3] = dh2_raw.sum(axis=1)
    dh2 = R(dh2_raw)
    g[P_W2:P_B2] = (dh2 @ h1.T).ravel()
    dh1_raw = W2b.T @ dh2
    mask1 = a1 > 0
    g[P_B1:P_W2] = np.where(mask1, dh1_raw, 0.0).sum(axis=1)
    dh1 = R(np.where(mask1, dh1_raw, 0.0))
    g[P_W1:P_B1] = (dh1 @ X).ravel()
    return loss, g, p


def scores(w, X):
    _, _, p = forward_backward(w, X, np.zeros(X.shape[0]))
    return p


def features(Qfilled: np.ndarray, cells_ij) -> np.ndarray:
    """Feature values (float64, before BF16 rounding) of the given (i, j)
    cells of a ghost-filled DG P=2 state Q[4][9][ny][nx]."""
    out = np.zeros((len(cells_ij), IN))
    offs = [(0, 0), (0, 1), (0, -1), (1, 0), (-1, 0)]
    for t, (i, j) in enumerate(cells_ij):
        mean = Qfilled[:, 0, i, j]
        r = abs(mean[0])
        scale = np.array([abs(mean[c]) + r + 1e-12 for c in range(4)])
        for n, (di, dj) in enumerate(offs):
            for c in range(4):
                for m in range(9):
                    v = Qfilled[c, m, i + di, j + dj]
                    if m == 0:
                        v = v - mean[c]
                    out[t, n * 36 + c * 9 + m] = np.float32(v / scale[c])
    return out
