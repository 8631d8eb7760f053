"""ORACLE — test infrastructure only. numpy restatement of the rank-predictor
MLP (8 -> 64 -> 64 -> 1, exact-erf GELU) with the library's parameter layout
and normalisation, its mean-squared-error loss and a hand-derived backward
pass. Pinned by central finite differences in tests/test_mlp_oracle.py.

Layout (4801 float32): W1[64][8] at 0, b1 at 512, W2[64][64] at 576
(W2[j][k]: hidden-1 unit k -> hidden-2 unit j), b2 at 4672, w3 at 4736, b3 at 4800.
norm (18): feature means[8], feature stds[8], target mean, target std.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

F, H = 8, 64
O_W1, O_B1 = 0, H * F
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + H
N_PARAMS = O_B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[O_W1:O_B1].reshape(H, F), p[O_B1:O_W2], p[O_W2:O_B2].reshape(H, H), p[O_B2:O_W3], p[O_W3:O_B3],
            p[O_B3])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, norm, X):
    """Returns predicted log-rank (de-normalised)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    xm, xs, ym, ys = norm[:F], norm[F:2 * F], norm[2 * F], norm[2 * F + 1]
    x = (np.asarray(X, np.float64) - xm) / xs
    h1 = gelu(x @ W1.T + b1)
    h2 = gelu(h1 @ W2.T + b2)
    return (h2 @ w3 + b3) * ys + ym


def loss(p, norm, X, y):
    ym, ys = norm[2 * F], norm[2 * F + 1]
    out_n = (forward(p, norm, X) - ym) / ys
    t = (np.asarray(y, np.float64) - ym) / ys
    return float(np.mean((out_n - t) ** 2))


def loss_grad(p, norm, X, y):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    xm, xs, ym, ys = norm[:F], norm[F:2 * F], norm[2 * F], norm[2 * F + 1]
    x = (np.asarray(X, np.float64) - xm) / xs
    t = (np.asarray(y, np.float64) - ym) / ys
    n = x.shape[0]
    z1 = x @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    o = h2 @ w3 + b3
    d = o - t
    dy = 2.0 * d / n
    g = np.zeros(N_PARAMS)
    g[O_B3] = dy.sum()
    g[O_W3:O_B3] = dy @ h2
    dz2 = np.outer(dy, w3) * gelu_d(z2)
    g[O_W2:O_B2] = (dz2.T @ h1).ravel()
    g[O_B2:O_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2) * gelu_d(z1)
    g[O_W1:O_B1] = (dz1.T @ x).ravel()
    g[O_B1:O_W2] = dz1.sum(0)
    return float(np.mean(d * d)), g
