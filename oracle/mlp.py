"""TEST INFRASTRUCTURE — double-precision restatement of the compensation
perceptron (10 -> 64 -> 64 -> 64 -> 1, ReLU), forward and backward.

Only tests/ may import this module, as the checker of the HIP kernels in
rexlib_amd/csrc/learned.hip. The reference has no such network (no
counterpart file); the restatement follows the layer algebra:
    h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), h3 = relu(h2 W3 + b3),
    out = h3 W4[:, 0] + b4[0], loss = mean((out - y)^2),
with the gradients of the chain rule. The kernels hold the weights and the
activations h1..h3 as float16 and accumulate in float32; `fp16_activations`
rounds the activations at the same points, so that the restatement and the
kernels differ only by accumulation order and the float16 gradients.
Pinned by tests/test_learned.py against central finite differences of the
loss, and against closed forms on a one-unit network.
"""
from __future__ import annotations

import numpy as np

SHAPES = [("w1", (16, 64)), ("w2", (64, 64)), ("w3", (64, 64)), ("w4", (64, 16)),
          ("b1", (64,)), ("b2", (64,)), ("b3", (64,)), ("b4", (16,))]


def unpack(flat):
    out = {}
    offset = 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = np.asarray(flat[offset:offset + size], dtype=np.float64).reshape(shape)
        offset += size
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def _round(a, fp16):
    return a.astype(np.float16).astype(np.float64) if fp16 else a


def forward(flat, x, fp16_activations=False):
    """out (M,) and the activations, for float64 parameters `flat` and inputs x (M, 16)."""
    p = unpack(flat)
    x = np.asarray(x, dtype=np.float64)
    h1 = _round(np.maximum(x @ p["w1"] + p["b1"], 0.0), fp16_activations)
    h2 = _round(np.maximum(h1 @ p["w2"] + p["b2"], 0.0), fp16_activations)
    h3 = _round(np.maximum(h2 @ p["w3"] + p["b3"], 0.0), fp16_activations)
    out = h3 @ p["w4"][:, 0] + p["b4"][0]
    return out, (x, h1, h2, h3)


def loss(flat, x, y, fp16_activations=False):
    out, _ = forward(flat, x, fp16_activations)
    return float(np.mean((out - np.asarray(y, dtype=np.float64)) ** 2))


def gradients(flat, x, y, fp16_activations=False):
    """(loss, gradient in the flat layout) of the mean squared error."""
    p = unpack(flat)
    out, (x, h1, h2, h3) = forward(flat, x, fp16_activations)
    m = x.shape[0]
    d_out = 2.0 * (out - np.asarray(y, dtype=np.float64)) / m
    g = {name: np.zeros(shape) for name, shape in SHAPES}
    g["w4"][:, 0] = h3.T @ d_out
    g["b4"][0] = d_out.sum()
    d3 = np.outer(d_out, p["w4"][:, 0]) * (h3 > 0)
    g["w3"] = h2.T @ d3
    g["b3"] = d3.sum(axis=0)
    d2 = (d3 @ p["w3"].T) * (h2 > 0)
    g["w2"] = h1.T @ d2
    g["b2"] = d2.sum(axis=0)
    d1 = (d2 @ p["w2"].T) * (h1 > 0)
    g["w1"] = x.T @ d1
    g["b1"] = d1.sum(axis=0)
    return float(np.mean((out - y) ** 2)), pack(g)
