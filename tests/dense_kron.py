"""Independent dense reference for tiny n (test helper).

Builds the full 2^n x 2^n circuit unitary from Kronecker products of the
textbook gate matrices, in PennyLane gate order, and the Hamiltonian from
Pauli matrices; wire w is bit w of the basis index, so a one-wire operator is
I^(n-1-w) (x) G (x) I^w.  Written from the definitions, not from the oracle.
"""

from __future__ import annotations

import numpy as np

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)


def ry(theta: float) -> np.ndarray:
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -s], [s, c]], dtype=complex)


def rz(phi: float) -> np.ndarray:
    return np.diag([np.exp(-0.5j * phi), np.exp(0.5j * phi)])


def on_wire(gate: np.ndarray, wire: int, n: int) -> np.ndarray:
    out = np.array([[1.0 + 0j]])
    for w in reversed(range(n)):
        out = np.kron(out, gate if w == wire else I2)
    return out


def cnot(control: int, target: int, n: int) -> np.ndarray:
    dim = 1 << n
    m = np.zeros((dim, dim), dtype=complex)
    for i in range(dim):
        j = i ^ (1 << target) if (i >> control) & 1 else i
        m[j, i] = 1.0
    return m


def unitary(n: int, params: np.ndarray) -> np.ndarray:
    u = np.eye(1 << n, dtype=complex)
    for layer in params:
        for w in range(n):
            u = on_wire(ry(layer[w, 0]), w, n) @ u
            u = on_wire(rz(layer[w, 1]), w, n) @ u
        for w in range(n):
            u = cnot(w, (w + 1) % n, n) @ u
    return u


def hamiltonian(n: int, field: float) -> np.ndarray:
    h = np.zeros((1 << n, 1 << n), dtype=complex)
    for w in range(n):
        h -= on_wire(Z, w, n) @ on_wire(Z, (w + 1) % n, n)
    for w in range(n):
        h -= field * on_wire(X, w, n)
    return h


def energy(n: int, field: float, params: np.ndarray) -> float:
    psi = unitary(n, params)[:, 0]
    return float(np.real(np.vdot(psi, hamiltonian(n, field) @ psi)))
