"""Warm-start MLP: the FP32 HIP kernels against the FP64 numpy oracle, and the
gain of the learned warm start over the random initialisation."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle

pytestmark = pytest.mark.gpu


def _net(backend_module, inputs, hidden, outputs, seed):
    return backend_module.WarmStartMLP(inputs, hidden, outputs, seed)


def test_forward_and_gradient_match_fp64_oracle(backend_module):
    inputs, hidden, outputs, rows = 3, 256, 80, 96
    net = _net(backend_module, inputs, hidden, outputs, 7)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 2, (rows, inputs)).astype(np.float32)
    t = rng.standard_normal((rows, outputs)).astype(np.float32)
    p = np.asarray(net.get_params(), dtype=np.float64)
    y = np.asarray(net.forward(x))
    y_ref, _ = mlp_oracle.forward(p, x, inputs, hidden, outputs)
    assert np.max(np.abs(y - y_ref)) <= 1e-5 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = net.loss_and_grad(x, t)
    loss_ref, g_ref = mlp_oracle.loss_and_grad(p, x, t, inputs, hidden, outputs)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    g = np.asarray(g, dtype=np.float64)
    print("max |dg| / max |g|", np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref)), flush=True)
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * np.max(np.abs(g_ref))


def test_training_steps_follow_oracle_adam(backend_module):
    inputs, hidden, outputs, rows, steps, lr = 3, 64, 12, 128, 8, 1e-3
    net = _net(backend_module, inputs, hidden, outputs, 9)
    rng = np.random.default_rng(10)
    x = rng.uniform(-1, 2, (rows, inputs)).astype(np.float32)
    t = rng.standard_normal((rows, outputs)).astype(np.float32)
    p0 = np.asarray(net.get_params(), dtype=np.float64)
    losses = np.asarray(net.train(x, t, steps, lr))
    p_ref, losses_ref = mlp_oracle.adam(p0, x, t, inputs, hidden, outputs, steps, lr)
    assert np.max(np.abs(losses - losses_ref)) <= 1e-4 * losses_ref[0]
    p = np.asarray(net.get_params(), dtype=np.float64)
    # Adam moves each weight by about lr per step; FP32 rounding stays far below that
    assert np.max(np.abs(p - p_ref)) <= 1e-2 * lr


@pytest.mark.parametrize("n,layers,count,mlp_steps,loss_drop", [(6, 3, 512, 1500, 0.1),
                                                                (10, 4, 1024, 2000, 0.2)])
def test_learned_warm_start_beats_random_init(n, layers, count, mlp_steps, loss_drop):
    """Plain, warm and branch-draw arms (the last starts every field from the
    selected training draw without the network); 10 qubits x 4 layers is the
    configured warm-start shape, here with 1024 training runs and 64 held-out
    fields instead of 4096 and 512."""
    from gpu_simulation_amd import warm_start as ws

    steps, lr = 300 if n == 10 else 200, 0.05
    print(f"case warm start n={n} L={layers} training runs={count}", flush=True)
    data = ws.generate_training_data(n, layers, count=count, steps=steps, lr=lr, seed=4)
    model = ws.WarmStartModel(n, layers, hidden=256, seed=0)
    losses = model.fit(data.fields, data.params, steps=mlp_steps, lr=1e-3)
    print("branch", data.branch, data.branch_mean_energies, "loss", losses[0], losses[-1],
          flush=True)
    # training makes progress (80 standardised outputs at 10 qubits x 4 layers
    # fit less closely than 36 at 6 x 3)
    assert losses[-1] < loss_drop * losses[0]
    branch_init = ws.reference_init(n, layers, ws.branch_seed(4, data.branch))
    res = ws.evaluate(model, np.random.default_rng(44).uniform(0.2, 2.0, 64), steps, lr, 1e-4,
                      branch_init=branch_init)
    print({k: v for k, v in res.items() if not isinstance(v, np.ndarray)}, flush=True)
    assert res["median_steps_warm"] < res["median_steps_plain"]
    # the warm runs converge at least as low as the plain ones on average
    assert res["mean_converged_energy_warm"] <= res["mean_converged_energy_plain"] + 1e-3
    # the network's predictions start lower than the shared draw of its branch
    assert res["mean_start_energy_warm"] < res["mean_start_energy_branch_draw"]
