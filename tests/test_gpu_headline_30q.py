"""GPU parity at the headline size: 30 qubits x 2 layers, complex128.

Three independent checks of the engine at the pass geometry of this size
(a 12-wire low tile and two full 9-wire high groups, wires 12..20 and
21..29, whose layer-1 chain segments are only built at n = 30):

* the CPU oracle's gate-level adjoint for the benchmark's own draw
  (``tests/golden/c5_30q2l.json``, made by ``tests/golden/make_c5_fixture.py``),
  at |dE|, |dgrad| <= 1e-10 under both backward schedules, with layer-1
  generation and the fused Hamiltonian each on and off;
* closed forms at full size: all angles zero leave |0...0>, so E = -n and
  dE/dtheta = -h [popcount(m) == 1] where m is e_w moved through the
  remaining ring layers, dE/dphi = 0; RY(pi/2) on every wire of layer 0 and
  every other angle zero gives |+>^n, so E = -n h;
* the exact parameter-shift identity dE/dp = (E(p + pi/2) - E(p - pi/2)) / 2
  (RY and RZ have generators +-1/2) on all 120 components, from energy-only
  evaluations, against the engine's adjoint gradient;
* a sparse brute force at full size: with only a few angles nonzero the state
  has a handful of nonzero amplitudes, so a dictionary simulation of the gate
  sequence (RY, RZ per wire, then CNOT(w, w+1) gate by gate) gives the exact
  energy and, by parameter shift, all 120 gradients, against the engine.
"""

from __future__ import annotations

import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest
from sparse_sim import ring_layer as _ring_layer
from sparse_sim import sparse_energy, sparse_grad

pytestmark = pytest.mark.gpu

TOL = 1e-10
N, LAYERS = 30, 2
FIXTURE = Path(__file__).resolve().parent / "golden" / "c5_30q2l.json"


def _fixture():
    data = json.loads(FIXTURE.read_text())
    params = np.random.default_rng(data["seed"]).uniform(
        0.0, 2.0 * math.pi, size=(data["layers"], data["num_qubits"], 2)).reshape(-1)
    digest = hashlib.sha256(np.ascontiguousarray(params, dtype="<f8").tobytes()).hexdigest()
    assert digest == data["params_sha256"], "regenerated parameters differ from the fixture's"
    return data, params


def test_fixture_parameters_and_bounds():
    data, params = _fixture()
    assert (data["num_qubits"], data["layers"], data["field"]) == (N, LAYERS, 1.0)
    assert len(data["gradient"]) == 2 * N * LAYERS == params.size
    assert abs(data["energy"]) <= N * (1.0 + abs(data["field"]))


@pytest.mark.parametrize("schedule", ["stored", "inverse_walk"])
def test_headline_against_oracle_fixture(backend_module, schedule):
    data, params = _fixture()
    g_ref = np.asarray(data["gradient"])
    eng = backend_module.RingIsingCudaBackend(N, LAYERS, data["field"], "structured_adjoint", 8,
                                              False)
    eng.set_schedule(schedule)
    outs = {}
    for gen in (1, 0):
        eng.set_layer1_generation(gen)
        for fused in (1, 0):
            eng.set_fused_hamiltonian(fused)
            print(f"case 30q x 2L schedule={schedule} generation={gen} fused={fused}", flush=True)
            out = eng.energy_and_grad(params)
            de = abs(out["energy"] - data["energy"])
            dg = float(np.max(np.abs(np.asarray(out["gradient"]) - g_ref)))
            print(f"  |dE| = {de:.3e}  max|dgrad| = {dg:.3e}", flush=True)
            assert de <= TOL, (gen, fused, de)
            assert dg <= TOL, (gen, fused, dg)
            outs[gen, fused] = out
        e_only = eng.energy_and_grad(params, False)
        assert abs(e_only["energy"] - data["energy"]) <= TOL
        assert np.asarray(e_only["gradient"]).size == 0


def test_headline_device_entry_point(backend_module):
    import torch

    data, params = _fixture()
    eng = backend_module.RingIsingCudaBackend(N, LAYERS, data["field"], "structured_adjoint", 8,
                                              False)
    p = torch.tensor(params, dtype=torch.float64, device="cuda")
    e = torch.empty(1, dtype=torch.float64, device="cuda")
    g = torch.empty_like(p)
    stream = torch.cuda.current_stream().cuda_stream
    print("case 30q x 2L device entry point", flush=True)
    eng.energy_and_grad_device(p.data_ptr(), e.data_ptr(), g.data_ptr(), stream)
    torch.cuda.synchronize()
    assert abs(float(e.item()) - data["energy"]) <= TOL
    assert float(np.max(np.abs(g.cpu().numpy() - np.asarray(data["gradient"])))) <= TOL


def test_headline_zero_angles_closed_form(backend_module):
    field = 0.7
    params = np.zeros(2 * N * LAYERS)
    eng = backend_module.RingIsingCudaBackend(N, LAYERS, field, "structured_adjoint", 8, False)
    print("case 30q x 2L zero angles", flush=True)
    out = eng.energy_and_grad(params)
    assert abs(out["energy"] + N) <= TOL
    grad = np.asarray(out["gradient"]).reshape(LAYERS, N, 2)
    assert np.max(np.abs(grad[:, :, 1])) <= TOL
    expected = np.zeros((LAYERS, N))
    for layer in range(LAYERS):
        for w in range(N):
            m = 1 << w
            for _ in range(LAYERS - layer):
                m = _ring_layer(m, N)
            expected[layer, w] = -field if bin(m).count("1") == 1 else 0.0
    # at 30 qubits one or two ring layers spread every single bit over two or
    # more wires, so the rule gives zero for every component (the nonzero
    # gradients are covered by the sparse brute force below)
    assert np.count_nonzero(expected) == 0
    assert np.max(np.abs(grad[:, :, 0] - expected)) <= TOL


def _sparse_point(seed):
    """Zero angles except six RY and four RZ angles spread over the low tile,
    both high groups and the wires whose pairs straddle them (0, 11, 12, 20,
    21, 29) in both layers."""
    rng = np.random.default_rng(seed)
    params = np.zeros((LAYERS, N, 2))
    for layer, w in [(0, 0), (0, 11), (0, 21), (1, 12), (1, 20), (1, 29)]:
        params[layer, w, 0] = rng.uniform(0.3, 2 * math.pi - 0.3)
    for layer, w in [(0, 12), (0, 29), (1, 0), (1, 21)]:
        params[layer, w, 1] = rng.uniform(0.3, 2 * math.pi - 0.3)
    return params.reshape(-1)


@pytest.mark.parametrize("schedule", ["stored", "inverse_walk"])
def test_headline_sparse_brute_force(backend_module, schedule):
    field = 0.9
    params = _sparse_point(11)
    e_ref = sparse_energy(N, LAYERS, field, params)
    g_ref = sparse_grad(N, LAYERS, field, params)
    # the point is not degenerate: the energy is not -n and many gradients are nonzero
    assert abs(e_ref + N) > 1e-3 and np.count_nonzero(np.abs(g_ref) > 1e-6) >= 10
    eng = backend_module.RingIsingCudaBackend(N, LAYERS, field, "structured_adjoint", 8, False)
    eng.set_schedule(schedule)
    for gen in (1, 0):
        eng.set_layer1_generation(gen)
        print(f"case 30q x 2L sparse point schedule={schedule} generation={gen}", flush=True)
        out = eng.energy_and_grad(params)
        de = abs(out["energy"] - e_ref)
        dg = float(np.max(np.abs(np.asarray(out["gradient"]) - g_ref)))
        print(f"  |dE| = {de:.3e}  max|dgrad| = {dg:.3e}", flush=True)
        assert de <= TOL and dg <= TOL


def test_headline_plus_state_energy(backend_module):
    field = 0.7
    params = np.zeros((LAYERS, N, 2))
    params[0, :, 0] = np.pi / 2
    eng = backend_module.RingIsingCudaBackend(N, LAYERS, field, "structured_adjoint", 8, False)
    print("case 30q x 2L plus state", flush=True)
    out = eng.energy_and_grad(params.reshape(-1))
    assert abs(out["energy"] + N * field) <= TOL
    e_only = eng.energy_and_grad(params.reshape(-1), False)
    assert abs(e_only["energy"] + N * field) <= TOL


def test_headline_parameter_shift_all_components(backend_module):
    data, params = _fixture()
    eng = backend_module.RingIsingCudaBackend(N, LAYERS, data["field"], "structured_adjoint", 8,
                                              False)
    print("case 30q x 2L parameter shift, 240 energy-only evaluations", flush=True)
    grad = np.asarray(eng.energy_and_grad(params)["gradient"])
    shift = np.empty_like(grad)
    for k in range(params.size):
        plus = params.copy()
        plus[k] += np.pi / 2
        minus = params.copy()
        minus[k] -= np.pi / 2
        e_plus = eng.energy_and_grad(plus, False)["energy"]
        e_minus = eng.energy_and_grad(minus, False)["energy"]
        shift[k] = 0.5 * (e_plus - e_minus)
    diff = float(np.max(np.abs(shift - grad)))
    print(f"  max|shift - adjoint| = {diff:.3e}", flush=True)
    assert diff <= TOL
    assert float(np.max(np.abs(shift - np.asarray(data["gradient"])))) <= TOL
