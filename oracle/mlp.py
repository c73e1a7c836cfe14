"""Float64 reference of the per-atom MLP used as the learned CG initial guess (test infrastructure).

The network (no counterpart in the reference; its specification is the build's own):
    h_0 = x (D0 inputs), h_l = silu(h_{l-1} W_l + b_l) for the hidden layers, y = h_L W_out + b_out,
    silu(z) = z / (1 + exp(-z)), loss = mean over atoms and outputs of (y - t)^2.
Forward, the analytic backward pass and an Adam step (Kingma and Ba, 2015: m = b1 m + (1-b1) g,
v = b2 v + (1-b2) g^2, w -= lr (m / (1-b1^s)) / (sqrt(v / (1-b2^s)) + eps)) are written out here so
that the GPU kernels can be checked against them.

Parity: pinned by central finite differences of the loss for the backward pass and by the
closed form of one Adam step (tests/test_oracle_mlp.py).
"""

from __future__ import annotations

import numpy as np


def silu(z: np.ndarray) -> np.ndarray:
    return z / (1.0 + np.exp(-z))


def dsilu(z: np.ndarray) -> np.ndarray:
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init(sizes: list[int], seed: int = 0) -> list[tuple[np.ndarray, np.ndarray]]:
    """Return [(W (in, out), b (out,))] with W ~ N(0, 1/in) (LeCun normal) and b = 0."""
    rng = np.random.default_rng(seed)
    return [(rng.normal(scale=1.0 / np.sqrt(a), size=(a, b)), np.zeros(b)) for a, b in zip(sizes[:-1], sizes[1:])]


def forward(params, x: np.ndarray) -> tuple[np.ndarray, list]:
    """Return (y, cache) for inputs x (B, D0); cache holds (h_in, z) per layer."""
    h = np.asarray(x, float)
    cache = []
    for li, (W, b) in enumerate(params):
        z = h @ W + b
        cache.append((h, z))
        h = z if li == len(params) - 1 else silu(z)
    return h, cache


def loss_grad(params, x: np.ndarray, t: np.ndarray) -> tuple[float, list]:
    """Return (loss, [(dW, db)]) of the mean squared error over all outputs."""
    y, cache = forward(params, x)
    B = y.size
    loss = float(np.sum((y - t) ** 2) / B)
    delta = 2.0 * (y - t) / B
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        h, z = cache[li]
        if li < len(params) - 1:
            delta = delta * dsilu(z)
        grads[li] = (h.T @ delta, delta.sum(0))
        delta = delta @ params[li][0].T
    return loss, grads


def adam_step(params, grads, state, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """Return (params, state) after one Adam step (step counts from 1); state = [(mW, vW, mb, vb)]."""
    out_p, out_s = [], []
    for (W, b), (gW, gb), (mW, vW, mb, vb) in zip(params, grads, state):
        mW = b1 * mW + (1 - b1) * gW
        vW = b2 * vW + (1 - b2) * gW * gW
        mb = b1 * mb + (1 - b1) * gb
        vb = b2 * vb + (1 - b2) * gb * gb
        c1, c2 = 1 - b1**step, 1 - b2**step
        W = W - lr * (mW / c1) / (np.sqrt(vW / c2) + eps)
        b = b - lr * (mb / c1) / (np.sqrt(vb / c2) + eps)
        out_p.append((W, b))
        out_s.append((mW, vW, mb, vb))
    return out_p, out_s


def zeros_state(params):
    return [(np.zeros_like(W), np.zeros_like(W), np.zeros_like(b), np.zeros_like(b)) for W, b in params]
