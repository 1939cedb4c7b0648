"""FP64 numpy restatement of the warm-start MLP (test infrastructure only).

Checker for the FP32 HIP kernels behind ``rie_mlp_*``: same packed parameter
layout (W1 [in x hidden], b1, W2 [hidden x hidden], b2, W3 [hidden x out], b3,
row-major, y = x W + b), exact GELU 0.5 x (1 + erf(x / sqrt 2)), mean squared
error over all rows x outputs values, and torch.optim.Adam's update.  The
MLP has no counterpart in the reference; these formulas are the textbook
back-propagation of the network, pinned in the tests by central finite
differences of this oracle's own loss.
"""

from __future__ import annotations

import math

import numpy as np

_erf = np.vectorize(math.erf, otypes=[np.float64])


def unpack(p: np.ndarray, inputs: int, hidden: int, outputs: int):
    p = np.asarray(p, dtype=np.float64)
    sizes = [inputs * hidden, hidden, hidden * hidden, hidden, hidden * outputs, outputs]
    parts, off = [], 0
    for s in sizes:
        parts.append(p[off:off + s])
        off += s
    if off != p.size:
        raise ValueError("parameter count does not match the sizes")
    w1, b1, w2, b2, w3, b3 = parts
    return (w1.reshape(inputs, hidden), b1, w2.reshape(hidden, hidden), b2,
            w3.reshape(hidden, outputs), b3)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def forward(p, x, inputs, hidden, outputs):
    w1, b1, w2, b2, w3, b3 = unpack(p, inputs, hidden, outputs)
    z1 = np.asarray(x, np.float64) @ w1 + b1
    a1 = gelu(z1)
    z2 = a1 @ w2 + b2
    a2 = gelu(z2)
    return a2 @ w3 + b3, (z1, a1, z2, a2)


def loss_and_grad(p, x, t, inputs, hidden, outputs):
    x = np.asarray(x, np.float64)
    y, (z1, a1, z2, a2) = forward(p, x, inputs, hidden, outputs)
    w1, b1, w2, b2, w3, b3 = unpack(p, inputs, hidden, outputs)
    d = y - np.asarray(t, np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / d.size
    g_w3 = a2.T @ dy
    g_b3 = dy.sum(axis=0)
    d2 = (dy @ w3.T) * gelu_grad(z2)
    g_w2 = a1.T @ d2
    g_b2 = d2.sum(axis=0)
    d1 = (d2 @ w2.T) * gelu_grad(z1)
    g_w1 = x.T @ d1
    g_b1 = d1.sum(axis=0)
    grad = np.concatenate([g_w1.ravel(), g_b1, g_w2.ravel(), g_b2, g_w3.ravel(), g_b3])
    return loss, grad


def adam(p, x, t, inputs, hidden, outputs, steps, lr, b1=0.9, b2=0.999, eps=1e-8):
    p = np.asarray(p, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for k in range(1, steps + 1):
        loss, g = loss_and_grad(p, x, t, inputs, hidden, outputs)
        losses.append(loss)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - (lr / (1 - b1 ** k)) * (m / (np.sqrt(v) / math.sqrt(1 - b2 ** k) + eps))
    return p, np.asarray(losses)
