"""CPU oracle of the surrogate MLP (numpy).

TEST INFRASTRUCTURE ONLY: checks the device forward/backward kernels of
libcac_hip.so (k_mlp.hip).  It evaluates the same network with the same
rounding points: inputs, weights, hidden activations and output/hidden
gradients rounded to BF16 (round to nearest even) before each product,
products and sums in float64 (the device accumulates in FP32).

``mode="bf16"`` mirrors the device; ``mode="fp64"`` drops every rounding, and
is what the finite-difference pin checks.
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(weights, x, mode="bf16"):
    """weights: list of (W [out][in], b [out]); returns y and the cache."""
    r = bf16 if mode == "bf16" else (lambda a: np.asarray(a, dtype=np.float64))
    h = r(np.asarray(x, dtype=np.float32) if mode == "bf16" else x)
    cache = [h]
    zs = []
    for l, (W, b) in enumerate(weights):
        z = h @ r(W).T + np.asarray(b, dtype=np.float64)
        if l < len(weights) - 1:
            zs.append(z)
            h = r(_silu(z).astype(np.float32)) if mode == "bf16" else _silu(z)
            cache.append(h)
        else:
            y = z
    return y, (cache, zs)


def loss_grads(weights, x, t, mode="bf16"):
    """MSE loss mean over n*out and its gradients [(dW, db)] per layer."""
    r = bf16 if mode == "bf16" else (lambda a: np.asarray(a, dtype=np.float64))
    y, (hs, zs) = forward(weights, x, mode)
    n, out = y.shape
    d = y - np.asarray(t, dtype=np.float64)
    loss = float((d * d).sum() / (n * out))
    dz = r((2.0 * d / (n * out)).astype(np.float32)) if mode == "bf16" else 2.0 * d / (n * out)
    grads = [None] * len(weights)
    for l in range(len(weights) - 1, -1, -1):
        W, _ = weights[l]
        grads[l] = (dz.T @ hs[l], dz.sum(axis=0))
        if l > 0:
            dh = dz @ r(W)
            g = dh * _dsilu(zs[l - 1])
            dz = r(g.astype(np.float32)) if mode == "bf16" else g
    return loss, grads
