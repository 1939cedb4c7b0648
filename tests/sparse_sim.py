"""Exact simulation of the ring-Ising circuit on a sparse (dictionary) state.

With only a few nonzero RY angles the state has a handful of nonzero
amplitudes at any qubit count, so this gives exact energies and gradients at
30 qubits and above.  The gate sequence is the gate-level one
(``ring_ising/models/ising.py:27-44``): per layer RY(theta_w) then RZ(phi_w) on
every wire, then CNOT(0, 1), CNOT(1, 2), ..., CNOT(n-1, 0) one gate at a time;
H = -sum_w Z_w Z_{w+1 mod n} - h sum_w X_w.  Gradients come from the exact
parameter-shift rule (generators +-1/2), not from an adjoint sweep, so the
check shares no algorithm with the engine or the oracle's adjoint.
"""

from __future__ import annotations

import math

import numpy as np


def ring_layer(index: int, n: int) -> int:
    """Basis index after CNOT(0,1), CNOT(1,2), ..., CNOT(n-1,0), gate by gate."""
    for w in range(n):
        if (index >> w) & 1:
            index ^= 1 << ((w + 1) % n)
    return index


def _ry(state: dict, bit: int, theta: float) -> dict:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    out: dict[int, complex] = {}
    for k, a in state.items():
        k0, k1 = k & ~bit, k | bit
        if k & bit:  # |1> -> -s|0> + c|1>
            out[k0] = out.get(k0, 0j) - s * a
            out[k1] = out.get(k1, 0j) + c * a
        else:  # |0> -> c|0> + s|1>
            out[k0] = out.get(k0, 0j) + c * a
            out[k1] = out.get(k1, 0j) + s * a
    return {k: a for k, a in out.items() if a != 0}


def _rz(state: dict, bit: int, phi: float) -> dict:
    lo = complex(math.cos(phi / 2), -math.sin(phi / 2))  # |0> -> e^{-i phi/2}
    hi = lo.conjugate()
    return {k: a * (hi if k & bit else lo) for k, a in state.items()}


def sparse_state(n: int, layers: int, params) -> dict:
    p = np.asarray(params, dtype=np.float64).reshape(layers, n, 2)
    state = {0: 1.0 + 0.0j}
    for layer in range(layers):
        for w in range(n):
            theta, phi = p[layer, w]
            if theta != 0.0:
                state = _ry(state, 1 << w, theta)
            if phi != 0.0:
                state = _rz(state, 1 << w, phi)
        state = {ring_layer(k, n): a for k, a in state.items()}
    return state


def sparse_energy(n: int, layers: int, field: float, params) -> float:
    state = sparse_state(n, layers, params)
    energy = 0.0
    for k, a in state.items():
        zz = sum(1 if ((k >> w) & 1) == ((k >> ((w + 1) % n)) & 1) else -1 for w in range(n))
        energy -= zz * abs(a) ** 2
        for w in range(n):
            b = state.get(k ^ (1 << w))
            if b is not None:
                energy -= field * (a.conjugate() * b).real
    return energy


def sparse_grad(n: int, layers: int, field: float, params) -> np.ndarray:
    flat = np.asarray(params, dtype=np.float64).reshape(-1)
    grad = np.empty_like(flat)
    for k in range(flat.size):
        plus, minus = flat.copy(), flat.copy()
        plus[k] += math.pi / 2
        minus[k] -= math.pi / 2
        grad[k] = 0.5 * (sparse_energy(n, layers, field, plus) -
                         sparse_energy(n, layers, field, minus))
    return grad
