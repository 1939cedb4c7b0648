"""torch.autograd wrapper: energy and gradient through the adjoint sweep.

Gradients are compared with the FP64 oracle (the PennyLane-order adjoint) at
<= 1e-10 and with central finite differences (step 1e-6; rounding limits that
comparison to ~1e-7).  Covers the chain-rule factor, interleaved forwards on
one engine, and the north-star sweep n = 8..24 at L = 2 plus 20q x 4L with and
without every-2-layer checkpoints.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_simulation_amd.autograd import RingIsingEnergy

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _params(seed, n, layers):
    return np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, size=(layers, n, 2))


def _energy_and_grad(module, host_params):
    p = torch.tensor(host_params, dtype=torch.float64, device=DEV, requires_grad=True)
    e = module(p)
    e.backward()
    return float(e.item()), p.grad.detach().cpu().numpy().reshape(-1)


@pytest.mark.parametrize("n", list(range(8, 25)))
def test_sweep_two_layers_against_oracle(oracle, n):
    params = _params(1000 + n, n, 2)
    print(f"case autograd n={n} L=2", flush=True)
    module = RingIsingEnergy(n, 2, 1.0)
    e, g = _energy_and_grad(module, params)
    e_ref, g_ref = oracle.energy_and_grad(n, 2, 1.0, params, form="structured")
    assert abs(e - e_ref) <= 1e-10
    assert np.max(np.abs(g - g_ref)) <= 1e-10


@pytest.mark.parametrize("every", [0, 2])
def test_twenty_qubits_four_layers(oracle, every):
    n, layers = 20, 4
    params = _params(2024, n, layers)
    module = RingIsingEnergy(n, layers, 1.0, checkpoint_every=every)
    e, g = _energy_and_grad(module, params)
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 1.0, params, form="structured")
    assert abs(e - e_ref) <= 1e-10
    assert np.max(np.abs(g - g_ref)) <= 1e-10
    drift = module.handle.native.inverse_walk_drift(params.reshape(-1))
    print(f"inverse-walk drift 20q x 4L, checkpoint_every={every}: {drift:.3e}")
    assert drift < 1e-11


def test_central_finite_differences():
    n, layers = 20, 4
    params = _params(7, n, layers)
    module = RingIsingEnergy(n, layers, 1.0)
    _, g = _energy_and_grad(module, params)
    flat = params.reshape(-1)
    rng = np.random.default_rng(0)
    step = 1e-6
    for k in rng.choice(flat.size, size=24, replace=False):
        with torch.no_grad():
            plus = flat.copy()
            plus[k] += step
            minus = flat.copy()
            minus[k] -= step
            ep = float(module(torch.tensor(plus.reshape(params.shape), device=DEV)).item())
            em = float(module(torch.tensor(minus.reshape(params.shape), device=DEV)).item())
        assert abs((ep - em) / (2 * step) - g[k]) < 2e-7, k


def test_chain_rule_scale_and_interleaved_forwards(oracle):
    for n in (10, 16):
        module = RingIsingEnergy(n, 3, 0.8)
        a = _params(1, n, 3)
        b = _params(2, n, 3)
        pa = torch.tensor(a, device=DEV, requires_grad=True)
        pb = torch.tensor(b, device=DEV, requires_grad=True)
        ea = module(pa)
        eb = module(pb)          # overwrites the resident state of the engine
        (3.0 * ea + 1.0).backward()
        (-0.5 * eb).backward()
        _, ga = oracle.energy_and_grad(n, 3, 0.8, a)
        _, gb = oracle.energy_and_grad(n, 3, 0.8, b)
        assert np.max(np.abs(pa.grad.cpu().numpy().reshape(-1) - 3.0 * ga)) <= 3e-10
        assert np.max(np.abs(pb.grad.cpu().numpy().reshape(-1) + 0.5 * gb)) <= 1e-10


def test_torch_optimizer_descends():
    module = RingIsingEnergy(12, 3, 1.0)
    p = torch.tensor(_params(4, 12, 3) * 0.1, device=DEV, requires_grad=True)
    opt = torch.optim.Adam([p], lr=0.05)
    first = None
    for _ in range(30):
        opt.zero_grad()
        e = module(p)
        e.backward()
        opt.step()
        first = first if first is not None else float(e.item())
    assert float(module(p).item()) < first


@pytest.mark.parametrize("tol, expect_active", [(0.0, True), (1.0, False)])
def test_drift_triggered_checkpoints(oracle, tol, expect_active):
    """drift_tol = 0 switches every-2-layer checkpoints on at the first call
    (any floating-point walk drifts above zero), drift_tol = 1 never does; the
    energy and gradient match the oracle either way, before and after."""
    n, layers = 20, 4
    params = _params(2024, n, layers)
    module = RingIsingEnergy(n, layers, 1.0, drift_tol=tol, drift_check_every=3)
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 1.0, params, form="structured")
    for call in range(4):
        print(f"case drift-triggered checkpoints tol={tol} call={call}", flush=True)
        e, g = _energy_and_grad(module, params)
        assert abs(e - e_ref) <= 1e-10
        assert np.max(np.abs(g - g_ref)) <= 1e-10
    assert module.last_drift is not None and 0.0 < module.last_drift < 1e-11
    assert module.checkpoints_active is expect_active
