"""CPU reference of the learned-cull MLP -- TEST INFRASTRUCTURE ONLY.

Only tests may import this module; the product never does. The network has
no counterpart in the reference: it is this project's 16-64-64-1 ReLU MLP
predicting the log of a projected Gaussian's largest alpha * T. This module
restates its forward pass and the gradient of its training loss in float64,
optionally rounding to bfloat16 (round to nearest even) exactly where the
gfx950 kernels feed the matrix cores:

  x   = bf16(float32((f - mu) * inv_sigma))
  a1  = bf16(relu(x W1b^T + b1))            W1b = bf16(W1)
  z2  = a1 W2b^T + b2,  a2 = relu(z2)       W2b = bf16(W2)
  y   = a2 . w3b + b3                       w3b = bf16(w3)
  loss = mean(w (y - t)^2), w = 1 + under_weight e^t if (t > log_thr and y < t) else 1

Backward: dy = 2 w (y - t) / n; dZ2 = dy w3b [z2 > 0]; dW2 = bf16(dZ2)^T a1;
dA1 = bf16(dZ2) W2b; dZ1 = dA1 [a1 > 0]; dW1 = bf16(dZ1)^T x; db1, db2 sum
the unrounded dZ1, dZ2; dw3 = sum dy a2; db3 = sum dy.

With rounding=False every bf16() is the identity and the gradient is the
exact derivative of the loss, which tests pin by finite differences.
"""
from __future__ import annotations

import numpy as np

IN, HIDDEN = 16, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + HIDDEN * IN
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + HIDDEN
PARAMS = OFF_B3 + 1  # 5313


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64 values."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(HIDDEN, IN), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(HIDDEN, HIDDEN), p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3],
            float(p[OFF_B3]))


def forward(p, mu, inv_sigma, feats, rounding=True) -> dict:
    rnd = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    f = np.asarray(feats, np.float32)
    if rounding:
        x = bf16((f - np.asarray(mu, np.float32)) * np.asarray(inv_sigma, np.float32))
    else:
        x = (f.astype(np.float64) - np.asarray(mu, np.float64)) * np.asarray(inv_sigma, np.float64)
    W1b, W2b, w3b = rnd(W1), rnd(W2), rnd(w3)
    z1 = x @ W1b.T + b1
    a1 = rnd(np.maximum(z1, 0.0))
    z2 = a1 @ W2b.T + b2
    a2 = np.maximum(z2, 0.0)
    y = a2 @ w3b + b3
    return {"x": x, "a1": a1, "z2": z2, "a2": a2, "y": y, "W2b": W2b, "w3b": w3b}


def loss_and_grad(p, mu, inv_sigma, feats, targets, under_weight, log_thr, rounding=True):
    rnd = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    fw = forward(p, mu, inv_sigma, feats, rounding)
    t = np.asarray(targets, np.float64)
    y = fw["y"]
    n = y.shape[0]
    d = y - t
    w = np.where((t > log_thr) & (y < t), 1.0 + under_weight * np.exp(t), 1.0)
    loss = float(np.mean(w * d * d))
    dy = 2.0 * w * d / n
    z2, a1, a2, x = fw["z2"], fw["a1"], fw["a2"], fw["x"]
    dz2 = dy[:, None] * fw["w3b"][None, :] * (z2 > 0)
    g = np.zeros(PARAMS, np.float64)
    g[OFF_W3:OFF_B3] = (dy[:, None] * a2).sum(0)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz2r = rnd(dz2)
    g[OFF_W2:OFF_B2] = (dz2r.T @ a1).ravel()
    da1 = dz2r @ fw["W2b"]
    dz1 = da1 * (a1 > 0)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (rnd(dz1).T @ x).ravel()
    g[OFF_B3] = dy.sum()
    return loss, g


def random_params(rng: np.random.Generator, scale: float = 1.0) -> np.ndarray:
    p = np.zeros(PARAMS, np.float32)
    p[OFF_W1:OFF_B1] = rng.uniform(-1, 1, HIDDEN * IN) * np.sqrt(6 / IN) * scale
    p[OFF_B1:OFF_W2] = rng.normal(0, 0.1, HIDDEN)
    p[OFF_W2:OFF_B2] = rng.uniform(-1, 1, HIDDEN * HIDDEN) * np.sqrt(6 / HIDDEN) * scale
    p[OFF_B2:OFF_W3] = rng.normal(0, 0.1, HIDDEN)
    p[OFF_W3:OFF_B3] = rng.uniform(-1, 1, HIDDEN) * np.sqrt(6 / HIDDEN)
    p[OFF_B3] = -8.0
    return p
