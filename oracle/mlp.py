"""TEST INFRASTRUCTURE ONLY — FP64 reference of the learned cost model (21 -> 64 -> 64 -> 1,
ReLU, MSE loss, Adam), with the parameter layout of the library's kernels:
W1[64][21] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.

No reference counterpart (the learned tuner extends the method). Pinned by central finite
differences of the loss (tests/test_oracle_mlp.py), independent of the analytic backward.
"""
from __future__ import annotations

import numpy as np

IN, H = 21, 64
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("w3", (H,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def init_params(seed: int) -> np.ndarray:
    """He-uniform initialisation (float32 values)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        fan_in = shape[1] if len(shape) == 2 else (IN if name == "b1" else H)
        if name.startswith("b"):
            parts.append(np.zeros(shape))
        else:
            lim = np.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, shape))
    return np.concatenate([a.ravel() for a in parts]).astype(np.float32)


def forward(p: np.ndarray, X: np.ndarray):
    q = unpack(p.astype(np.float64))
    X = X.astype(np.float64)
    z1 = X @ q["W1"].T + q["b1"]
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ q["W2"].T + q["b2"]
    a2 = np.maximum(z2, 0.0)
    y = a2 @ q["w3"] + q["b3"][0]
    return z1, z2, y


def loss(p: np.ndarray, X: np.ndarray, T: np.ndarray) -> float:
    _, _, y = forward(p, X)
    return float(np.mean((y - T.astype(np.float64)) ** 2))


def loss_grad(p: np.ndarray, X: np.ndarray, T: np.ndarray):
    """MSE loss and its analytic gradient, FP64."""
    q = unpack(p.astype(np.float64))
    X = X.astype(np.float64)
    B = X.shape[0]
    z1, z2, y = forward(p, X)
    a1, a2 = np.maximum(z1, 0.0), np.maximum(z2, 0.0)
    dy = 2.0 * (y - T.astype(np.float64)) / B
    dz2 = np.outer(dy, q["w3"]) * (z2 > 0)
    dz1 = (dz2 @ q["W2"]) * (z1 > 0)
    g = {"W1": dz1.T @ X, "b1": dz1.sum(0), "W2": dz2.T @ a1, "b2": dz2.sum(0), "w3": a2.T @ dy,
         "b3": np.array([dy.sum()])}
    return float(np.mean((y - T) ** 2)), np.concatenate([g[n].ravel() for n, _ in SIZES])


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** t, 1 - b2 ** t
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def features(n, c, h, w, k, cfg) -> np.ndarray:
    f = np.zeros(IN, dtype=np.float32)
    f[:5] = np.log2([n, c, h, w, k])
    f[5 + cfg] = 1.0
    return f
