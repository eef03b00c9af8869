"""FP64 CPU reference of the learned Riemann-guess network (test infrastructure).

The learned starting guess has no counterpart in the reference, so this file
is the checker for the GPU training and inference kernels: the same features,
network and mean-squared loss in FP64 numpy, with the backward pass written
out by hand and itself pinned by central finite differences
(tests/test_mlp_oracle.py).

Network 6 -> 64 -> 64 -> 64 -> 1 with ReLU; parameters in one vector
[W1 (64x6), b1, W2 (64x64), b2, W3 (64x64), b3, w4, b4] (8833 values).  Features of a pair row
(u_L, rho_L, p_L, u_R, rho_R, p_R, p*): with pair means rho_m, p_m, u_m and
c_m = sqrt(gamma p_m / rho_m): (rho/rho_m, p/p_m, (u-u_m)/c_m) for L and R;
target log(p*/p_m).
"""
from __future__ import annotations

import numpy as np

IN, H = 6, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + H
N_PARAMS = OFF_B4 + 1


def unpack(w: np.ndarray):
    w = np.asarray(w, dtype=np.float64)
    W1 = w[OFF_W1:OFF_B1].reshape(H, IN)
    b1 = w[OFF_B1:OFF_W2]
    W2 = w[OFF_W2:OFF_B2].reshape(H, H)
    b2 = w[OFF_B2:OFF_W3]
    W3 = w[OFF_W3:OFF_B3].reshape(H, H)
    b3 = w[OFF_B3:OFF_W4]
    w4 = w[OFF_W4:OFF_B4]
    b4 = w[OFF_B4]
    return W1, b1, W2, b2, W3, b3, w4, b4


def features(rows: np.ndarray, gamma: float):
    r = np.asarray(rows, dtype=np.float64)
    uL, rL, pL, uR, rR, pR = (r[:, k] for k in range(6))
    r_m = 0.5 * (rL + rR)
    p_m = 0.5 * (pL + pR)
    u_m = 0.5 * (uL + uR)
    c_m = np.sqrt(gamma * p_m / r_m)
    x = np.stack([rL / r_m, pL / p_m, (uL - u_m) / c_m, rR / r_m, pR / p_m, (uR - u_m) / c_m], axis=1)
    return x, p_m


def forward(w: np.ndarray, x: np.ndarray):
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(w)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ W3.T + b3
    h3 = np.maximum(z3, 0.0)
    out = h3 @ w4 + b4
    return out, (z1, h1, z2, h2, z3, h3)


def loss_grad(w: np.ndarray, rows: np.ndarray, gamma: float):
    """Mean squared error of log(p*/p_m) and its gradient (FP64)."""
    x, p_m = features(rows, gamma)
    y = np.log(np.asarray(rows, dtype=np.float64)[:, 6] / p_m)
    out, (z1, h1, z2, h2, z3, h3) = forward(w, x)
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(w)
    B = len(y)
    d = out - y
    loss = float(np.mean(d * d))
    dout = 2.0 * d / B
    g = np.zeros(N_PARAMS)
    g[OFF_W4:OFF_B4] = h3.T @ dout
    g[OFF_B4] = dout.sum()
    dz3 = np.outer(dout, w4) * (z3 > 0)
    g[OFF_W3:OFF_B3] = (dz3.T @ h2).ravel()
    g[OFF_B3:OFF_W4] = dz3.sum(axis=0)
    dz2 = (dz3 @ W3) * (z2 > 0)
    g[OFF_W2:OFF_B2] = (dz2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    g[OFF_W1:OFF_B1] = (dz1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    return loss, g


def predict(w: np.ndarray, rows: np.ndarray, gamma: float) -> np.ndarray:
    x, p_m = features(rows, gamma)
    out, _ = forward(w, x)
    return p_m * np.exp(out)
