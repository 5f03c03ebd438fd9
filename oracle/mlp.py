"""numpy restatement of the stencil MLP (test infrastructure only).

Forward: h1 = tanh(x W1 + b1), h2 = tanh(h1 W2 + b2), y = h2 W3 + b3.
Loss: mean over all outputs of (y - t)^2.  Backward by hand (the GPU kernels'
gradients are compared with these; this restatement is itself pinned by a
central finite-difference check in tests/test_oracle.py).  BF16 inference is
modelled by rounding inputs, weights and hidden activations to bfloat16
(round-to-nearest-even) and accumulating in float64.
"""
from __future__ import annotations

import numpy as np


def forward(x, w1, b1, w2, b2, w3, b3):
    x = x.astype(np.float64)
    h1 = np.tanh(x @ w1 + b1)
    h2 = np.tanh(h1 @ w2 + b2)
    y = h2 @ w3 + b3
    return y, (x, h1, h2)


def loss_and_grads(x, t, w1, b1, w2, b2, w3, b3):
    y, (x, h1, h2) = forward(x, w1, b1, w2, b2, w3, b3)
    n = y.size
    e = y - t
    loss = float((e * e).sum() / n)
    d3 = 2.0 * e / n
    g_w3 = h2.T @ d3
    g_b3 = d3.sum(0)
    d2 = (d3 @ w3.T) * (1.0 - h2 * h2)
    g_w2 = h1.T @ d2
    g_b2 = d2.sum(0)
    d1 = (d2 @ w2.T) * (1.0 - h1 * h1)
    g_w1 = x.T @ d1
    g_b1 = d1.sum(0)
    return y, loss, [g_w1, g_b1, g_w2, g_b2, g_w3, g_b3]


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def infer_bf16(x, w1, b1, w2, b2, w3, b3):
    h1 = bf16(np.tanh(bf16(x) @ bf16(w1) + b1).astype(np.float32))
    h2 = bf16(np.tanh(h1 @ bf16(w2) + b2).astype(np.float32))
    return h2 @ bf16(w3) + b3
