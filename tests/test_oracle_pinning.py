"""Checks that pin each oracle function independently of its own formulas.

* closed forms: theta = phi = 0 -> |0...0>, E = -n, zero gradient;
  RY(pi/2) everywhere -> |+>^n, invariant under CNOT, E = -n h;
* exact parameter-shift rule for RY/RZ (generators +-1/2) vs the adjoint
  gradient, for both restatements; central finite differences;
* an independent dense Kronecker simulation (tests/dense_kron.py) for the
  state, H psi and the energy, n <= 6 (covers the doubled ZZ term at n = 2);
* the closed-form ring permutation vs n sequential CNOTs, and its inverse;
* gate-level restatement == structured restatement; inverse-walk drift.
A plausible mistake (dropped term, wrong sign or index, transposed factor)
breaks at least one of these; the last test shows the gradient checks catch a
one-parameter sign flip planted in the oracle's output.
"""

from __future__ import annotations

import numpy as np
import pytest

import dense_kron


def _params(seed, n, layers, lo=0.0, hi=2 * np.pi):
    return np.random.default_rng(seed).uniform(lo, hi, size=(layers, n, 2))


def _ring_layer(index: int, n: int) -> int:
    """Basis index after CNOT(0,1), CNOT(1,2), ..., CNOT(n-1,0), gate by gate."""
    for w in range(n):
        if (index >> w) & 1:
            index ^= 1 << ((w + 1) % n)
    return index


@pytest.mark.parametrize("form", ["gate", "structured"])
@pytest.mark.parametrize("n,layers", [(2, 2), (3, 3), (5, 2), (9, 3)])
def test_zero_angles_closed_form(oracle, form, n, layers):
    # All angles zero: the state stays |0...0>, E = -n.  A small RY(d) on wire
    # w of layer l adds (d/2)|e_w>, which the remaining L - l ring layers move
    # to |m>; hence dE/dtheta = 2 Re <0|H|m> / 2 = -h [popcount(m) == 1] and
    # every dE/dphi = 0 (RZ only phases |0>).
    field = 0.7
    params = np.zeros((layers, n, 2))
    energy, grad = oracle.energy_and_grad(n, layers, field, params, form=form)
    assert abs(energy + n) < 1e-14
    grad = grad.reshape(layers, n, 2)
    assert np.max(np.abs(grad[:, :, 1])) < 1e-14
    for l in range(layers):
        for w in range(n):
            m = 1 << w
            for _ in range(layers - l):
                m = _ring_layer(m, n)
            expected = -field if bin(m).count("1") == 1 else 0.0
            assert abs(grad[l, w, 0] - expected) < 1e-14, (l, w)


@pytest.mark.parametrize("form", ["gate", "structured"])
@pytest.mark.parametrize("n,field", [(2, 0.3), (4, 1.0), (7, 1.7)])
def test_plus_state_energy(oracle, form, n, field):
    # RY(pi/2) on every wire in layer 0, every other angle zero: |+>^n, which
    # the ring CNOTs and identity rotations leave unchanged
    params = np.zeros((3, n, 2))
    params[0, :, 0] = np.pi / 2
    energy, _ = oracle.energy_and_grad(n, 3, field, params, form=form)
    assert abs(energy + n * field) < 1e-13


@pytest.mark.parametrize("form", ["gate", "structured"])
@pytest.mark.parametrize("n,layers", [(2, 2), (3, 1), (4, 3), (6, 2)])
def test_parameter_shift_rule(oracle, form, n, layers):
    params = _params(10 * n + layers, n, layers)
    _, grad = oracle.energy_and_grad(n, layers, 0.9, params, form=form)
    shift = oracle.param_shift_grad(n, layers, 0.9, params, form=form)
    assert np.max(np.abs(grad - shift)) < 1e-12


@pytest.mark.parametrize("n,layers", [(3, 2), (5, 2)])
def test_central_finite_differences(oracle, n, layers):
    params = _params(7 + n, n, layers)
    _, grad = oracle.energy_and_grad(n, layers, 1.1, params)
    fd = oracle.central_fd_grad(n, layers, 1.1, params, step=1e-6)
    assert np.max(np.abs(grad - fd)) < 1e-8


@pytest.mark.parametrize("n,layers,field", [(2, 2, 0.7), (3, 2, 1.0), (4, 1, 1.3), (6, 2, 1.1)])
def test_dense_kronecker_reference(oracle, n, layers, field):
    params = _params(100 + n, n, layers)
    psi = oracle.state(n, layers, params)
    ref_psi = dense_kron.unitary(n, params)[:, 0]
    assert np.max(np.abs(psi - ref_psi)) < 1e-13
    h_psi = oracle.apply_hamiltonian(n, field, psi)
    assert np.max(np.abs(h_psi - dense_kron.hamiltonian(n, field) @ psi)) < 1e-12
    energy, _ = oracle.energy_and_grad(n, layers, field, params)
    assert abs(energy - dense_kron.energy(n, field, params)) < 1e-12


def test_two_qubit_ring_doubles_zz(oracle):
    # n = 2: (0,1) and (1,0) are both ring edges -> H = -2 Z0 Z1 - h (X0 + X1)
    params = _params(5, 2, 2)
    psi = dense_kron.unitary(2, params)[:, 0]
    h = -2 * np.kron(dense_kron.Z, dense_kron.Z) - 0.7 * (
        np.kron(dense_kron.X, dense_kron.I2) + np.kron(dense_kron.I2, dense_kron.X))
    energy, _ = oracle.energy_and_grad(2, 2, 0.7, params)
    assert abs(energy - np.real(np.vdot(psi, h @ psi))) < 1e-13


@pytest.mark.parametrize("n", [2, 3, 5, 8, 10])
def test_ring_permutation_closed_form(oracle, n):
    dim = 1 << n
    for i in range(dim):
        # apply CNOT(0,1), CNOT(1,2), ..., CNOT(n-1,0) to the basis index
        j = i
        for w in range(n):
            if (j >> w) & 1:
                j ^= 1 << ((w + 1) % n)
        assert oracle.ring_cnot_index(i, n, inverse=False) == j
        assert oracle.ring_cnot_index(j, n, inverse=True) == i


@pytest.mark.parametrize("n,layers", [(2, 3), (5, 2), (8, 3), (11, 2), (14, 2)])
def test_gate_level_equals_structured(oracle, n, layers):
    params = _params(n * layers, n, layers)
    e_gate, g_gate = oracle.energy_and_grad(n, layers, 0.8, params, form="gate")
    e_str, g_str = oracle.energy_and_grad(n, layers, 0.8, params, form="structured")
    assert abs(e_gate - e_str) < 1e-12
    assert np.max(np.abs(g_gate - g_str)) < 1e-12


def test_inverse_walk_returns_to_zero_state(oracle):
    params = _params(3, 10, 6)
    assert oracle.structured_drift(10, 6, 1.0, params) < 1e-12


def test_energy_is_bounded(oracle):
    for seed in range(5):
        params = _params(seed, 7, 3)
        energy, _ = oracle.energy_and_grad(7, 3, 1.5, params)
        assert abs(energy) <= 7 * (1 + 1.5) + 1e-12


def test_gradient_checks_catch_a_planted_sign_error(oracle):
    params = _params(77, 4, 2)
    _, grad = oracle.energy_and_grad(4, 2, 1.0, params)
    planted = grad.copy()
    k = int(np.argmax(np.abs(grad)))
    planted[k] = -planted[k]  # a sign mistake on one parameter
    shift = oracle.param_shift_grad(4, 2, 1.0, params)
    assert np.max(np.abs(grad - shift)) < 1e-12
    assert np.max(np.abs(planted - shift)) > 1e-3


@pytest.mark.parametrize("n,layers", [(6, 3), (10, 2), (13, 2)])
def test_sparse_simulation_matches_the_oracle(oracle, n, layers):
    # the dictionary simulation used at 30 qubits (tests/sparse_sim.py), checked
    # against both oracle forms where a dense state fits: energy and every
    # gradient component at a point with a few nonzero angles
    from sparse_sim import sparse_energy, sparse_grad

    rng = np.random.default_rng(17 + n)
    field = 0.8
    params = np.zeros((layers, n, 2))
    params[:, ::3, 0] = rng.uniform(0.0, 2 * np.pi, params[:, ::3, 0].shape)
    params[:, 1::4, 1] = rng.uniform(0.0, 2 * np.pi, params[:, 1::4, 1].shape)
    e = sparse_energy(n, layers, field, params)
    g = sparse_grad(n, layers, field, params)
    assert np.count_nonzero(np.abs(g) > 1e-6) >= 4
    for form in ("gate", "structured"):
        e_ref, g_ref = oracle.energy_and_grad(n, layers, field, params, form=form)
        assert abs(e - e_ref) <= 1e-12
        assert np.max(np.abs(g - g_ref)) <= 1e-12
