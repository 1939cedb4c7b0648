"""The warm-start MLP oracle's back-propagation against central finite
differences of its own loss (CPU)."""

from __future__ import annotations

import numpy as np

from oracle import mlp_oracle


def test_backprop_matches_central_differences():
    inputs, hidden, outputs, rows = 3, 5, 4, 7
    rng = np.random.default_rng(0)
    count = inputs * hidden + hidden + hidden * hidden + hidden + hidden * outputs + outputs
    p = rng.uniform(-0.8, 0.8, count)
    x = rng.uniform(-1, 2, (rows, inputs))
    t = rng.standard_normal((rows, outputs))
    _, g = mlp_oracle.loss_and_grad(p, x, t, inputs, hidden, outputs)
    eps = 1e-6
    fd = np.empty(count)
    for i in range(count):
        e = np.zeros(count)
        e[i] = eps
        lp, _ = mlp_oracle.loss_and_grad(p + e, x, t, inputs, hidden, outputs)
        lm, _ = mlp_oracle.loss_and_grad(p - e, x, t, inputs, hidden, outputs)
        fd[i] = (lp - lm) / (2 * eps)
    assert np.max(np.abs(fd - g)) <= 1e-8


def test_gelu_reference_values():
    # GELU(x) = x * Phi(x): Phi(0) = 1/2, Phi(1) = 0.8413447460685429...
    assert mlp_oracle.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mlp_oracle.gelu(np.array([1.0]))[0] - 0.8413447460685429) <= 1e-15
    assert abs(mlp_oracle.gelu(np.array([-1.0]))[0] + 0.15865525393145707) <= 1e-15
    x = np.linspace(-3, 3, 13)
    h = 1e-6
    fd = (mlp_oracle.gelu(x + h) - mlp_oracle.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp_oracle.gelu_grad(x))) <= 1e-9


def test_adam_first_step_closed_form():
    # after one step m_hat = g and v_hat = g^2 exactly, so
    # p1 = p0 - lr * g / (|g| + eps): a sign step of size lr where |g| >> eps
    inputs, hidden, outputs = 3, 8, 5
    rng = np.random.default_rng(3)
    count = inputs * hidden + hidden + hidden * hidden + hidden + hidden * outputs + outputs
    p0 = 0.3 * rng.standard_normal(count)
    x = rng.uniform(-1, 1, (16, inputs))
    t = rng.standard_normal((16, outputs))
    lr, eps = 1e-2, 1e-8
    _, g = mlp_oracle.loss_and_grad(p0, x, t, inputs, hidden, outputs)
    p1, losses = mlp_oracle.adam(p0, x, t, inputs, hidden, outputs, 1, lr, eps=eps)
    np.testing.assert_allclose(p1, p0 - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)
    assert losses.shape == (1,)
