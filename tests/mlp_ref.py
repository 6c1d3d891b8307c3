"""float64 numpy restatement of the scheduling network (8 -> 64 -> 64 -> 2, ReLU) and of its
loss gradient: the checker for the device kernels in poa_amd/csrc/mlp.hip."""
from __future__ import annotations

import numpy as np

IN, H, OUT = 8, 64, 2
SHAPES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, OUT)),
          ("b3", (OUT,))]
N_PARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w16: np.ndarray) -> dict:
    """FP16 bit patterns to float64 arrays (the values the device computes with)."""
    f = np.asarray(w16, np.uint16).view(np.float16).astype(np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = f[o:o + n].reshape(shape)
        o += n
    return out


def forward(P: dict, X: np.ndarray):
    a1 = X @ P["W1"] + P["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ P["W2"] + P["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ P["W3"] + P["b3"]
    return y, (h1, h2)


def loss_and_grad(P: dict, X: np.ndarray, T: np.ndarray):
    """L = 1/(2B) sum (y - t)^2 and dL/dparams, flattened in the device's order."""
    B = len(X)
    y, (h1, h2) = forward(P, X)
    d = y - T
    L = 0.5 * np.sum(d * d) / B
    dy = d / B
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    d2 = (dy @ P["W3"].T) * (h2 > 0)
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = (d2 @ P["W2"].T) * (h1 > 0)
    g["W1"] = X.T @ d1
    g["b1"] = d1.sum(0)
    flat = np.concatenate([g[name].ravel() for name, _ in SHAPES])
    return L, flat
