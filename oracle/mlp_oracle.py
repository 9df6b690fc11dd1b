"""CPU oracle of the cost model (float64 numpy).  TEST INFRASTRUCTURE ONLY:
used by the tests to check the HIP forward pass and gradients.

The network (no counterpart in the reference): standardised input x (8) ->
relu(x W1 + b1) (64) -> relu(. W2 + b2) (64) -> . w3 + b3; loss = mean (y - t)^2.
Gradients are checked by central finite differences in float64, which pins the
analytic backward pass independently of any formula restated here.
"""
from __future__ import annotations

import numpy as np


def forward(p: dict, X: np.ndarray) -> np.ndarray:
    x = (np.asarray(X, np.float64) - p["mu"].astype(np.float64)) / p["sd"].astype(np.float64)
    a1 = np.maximum(x @ p["w1"].astype(np.float64) + p["b1"].astype(np.float64), 0.0)
    a2 = np.maximum(a1 @ p["w2"].astype(np.float64) + p["b2"].astype(np.float64), 0.0)
    return a2 @ p["w3"].astype(np.float64) + float(p["b3"])


def loss(p: dict, X: np.ndarray, t: np.ndarray) -> float:
    y = forward(p, X)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def finite_difference(p: dict, X, t, name: str, index: tuple, h: float = 1e-4) -> float:
    """d loss / d p[name][index] by central differences in float64."""
    q = {k: np.array(v, dtype=np.float64, copy=True) for k, v in p.items()}
    base = q[name][index] if q[name].ndim else float(q[name])
    if q[name].ndim:
        q[name][index] = base + h
        lp = loss(q, X, t)
        q[name][index] = base - h
        lm = loss(q, X, t)
    else:
        q[name] = np.float64(base + h)
        lp = loss(q, X, t)
        q[name] = np.float64(base - h)
        lm = loss(q, X, t)
    return (lp - lm) / (2 * h)
