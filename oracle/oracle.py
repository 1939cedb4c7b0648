"""ctypes front end of the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module.  It wraps
``oracle/ring_ising_oracle.c`` (built into ``oracle/lib/`` by
``oracle/Makefile``) and restates, in numpy, the reference's
parameter initialisation and gradient-descent loop so the golden energies of
the reference's 20-step runs can be reproduced end to end:

* ``make_initial_params`` follows ``ring_ising/params.py:8-20``
  (``init_scale * default_rng(seed).standard_normal((L, n, 2))``);
* ``gd_loop`` follows ``ring_ising/training/loop.py:93-106`` and
  ``ring_ising/workflows/common.py:23-25`` (``p - lr * g`` in float64), with
  the final energy evaluated on the final parameters
  (``ring_ising/workflows/standalone.py:133``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libring_ising_oracle.so"

_lib = None


def build(force: bool = False) -> Path:
    """Compile the oracle with ``make`` (gcc + OpenMP)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(LIB_PATH))
    dp = ctypes.POINTER(ctypes.c_double)
    lib.rio_gate_energy_and_grad.argtypes = [
        ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int, dp, dp]
    lib.rio_gate_energy_and_grad.restype = ctypes.c_int
    lib.rio_structured_energy_and_grad.argtypes = [
        ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int, dp, dp, dp]
    lib.rio_structured_energy_and_grad.restype = ctypes.c_int
    lib.rio_gate_state.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
    lib.rio_gate_state.restype = ctypes.c_int
    lib.rio_apply_hamiltonian.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
    lib.rio_apply_hamiltonian.restype = ctypes.c_int
    lib.rio_ring_cnot_index.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    lib.rio_ring_cnot_index.restype = ctypes.c_uint64
    lib.rio_set_num_threads.argtypes = [ctypes.c_int]
    lib.rio_set_num_threads.restype = None
    lib.rio_get_max_threads.argtypes = []
    lib.rio_get_max_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _dptr(array: np.ndarray):
    return array.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _flat_params(num_qubits: int, layers: int, params) -> np.ndarray:
    flat = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
    if flat.size != 2 * num_qubits * layers:
        raise ValueError("params does not match num_qubits/layers.")
    return flat


def set_num_threads(threads: int) -> None:
    _load().rio_set_num_threads(int(threads))


def max_threads() -> int:
    return int(_load().rio_get_max_threads())


def energy_and_grad(num_qubits: int, layers: int, field: float, params,
                    *, compute_gradient: bool = True, form: str = "gate"):
    """Energy and flat gradient ``[(l*n + w)*2 + {0: theta_RY, 1: phi_RZ}]``.

    ``form="gate"`` is the PennyLane-order gate-level adjoint; ``form="structured"``
    is the fused/closed-form restatement.  Returns ``(energy, grad)``; ``grad``
    is ``None`` when ``compute_gradient`` is false.
    """
    lib = _load()
    flat = _flat_params(num_qubits, layers, params)
    energy = np.zeros(1)
    grad = np.zeros(flat.size)
    if form == "gate":
        status = lib.rio_gate_energy_and_grad(
            num_qubits, layers, float(field), _dptr(flat), int(compute_gradient),
            _dptr(energy), _dptr(grad))
    elif form == "structured":
        drift = np.zeros(1)
        status = lib.rio_structured_energy_and_grad(
            num_qubits, layers, float(field), _dptr(flat), int(compute_gradient),
            _dptr(energy), _dptr(grad), _dptr(drift))
    else:
        raise ValueError(f"unknown oracle form {form!r}")
    if status != 0:
        raise RuntimeError(f"oracle failed with status {status}")
    return float(energy[0]), (grad if compute_gradient else None)


def structured_drift(num_qubits: int, layers: int, field: float, params) -> float:
    """|| current - |0> || after the oracle's full inverse walk."""
    lib = _load()
    flat = _flat_params(num_qubits, layers, params)
    energy = np.zeros(1)
    grad = np.zeros(flat.size)
    drift = np.zeros(1)
    status = lib.rio_structured_energy_and_grad(
        num_qubits, layers, float(field), _dptr(flat), 1, _dptr(energy), _dptr(grad),
        _dptr(drift))
    if status != 0:
        raise RuntimeError(f"oracle failed with status {status}")
    return float(drift[0])


def state(num_qubits: int, layers: int, params) -> np.ndarray:
    """Final statevector (complex128, index bit w = wire w)."""
    lib = _load()
    flat = _flat_params(num_qubits, layers, params)
    out = np.zeros(2 << num_qubits)
    if lib.rio_gate_state(num_qubits, layers, _dptr(flat), _dptr(out)) != 0:
        raise RuntimeError("oracle state failed")
    return out.view(np.complex128)


def apply_hamiltonian(num_qubits: int, field: float, vec: np.ndarray) -> np.ndarray:
    lib = _load()
    src = np.ascontiguousarray(np.asarray(vec, dtype=np.complex128)).view(np.float64)
    out = np.zeros_like(src)
    if lib.rio_apply_hamiltonian(num_qubits, float(field), _dptr(src), _dptr(out)) != 0:
        raise RuntimeError("oracle hamiltonian failed")
    return out.view(np.complex128)


def ring_cnot_index(index: int, num_qubits: int, inverse: bool = False) -> int:
    return int(_load().rio_ring_cnot_index(int(index), int(num_qubits), int(inverse)))


def make_initial_params(num_qubits: int, layers: int, seed: int, init_scale: float) -> np.ndarray:
    """ring_ising/params.py:8-20."""
    rng = np.random.default_rng(seed)
    return np.asarray(init_scale * rng.standard_normal((layers, num_qubits, 2)),
                      dtype=np.float64)


def gd_loop(num_qubits: int, layers: int, field: float, params, steps: int, stepsize: float,
            *, form: str = "gate"):
    """Plain gradient descent; returns (final_params, step_energies, final_energy)."""
    current = np.asarray(params, dtype=np.float64).reshape(layers, num_qubits, 2)
    energies = []
    for _ in range(steps):
        energy, grad = energy_and_grad(num_qubits, layers, field, current, form=form)
        energies.append(energy)
        current = current - stepsize * grad.reshape(current.shape)
    final_energy, _ = energy_and_grad(num_qubits, layers, field, current,
                                      compute_gradient=False, form=form)
    return current, energies, final_energy


def param_shift_grad(num_qubits: int, layers: int, field: float, params,
                     *, form: str = "gate") -> np.ndarray:
    """dE/dp = (E(p + pi/2) - E(p - pi/2)) / 2, exact for RY/RZ (generators +-1/2)."""
    flat = _flat_params(num_qubits, layers, params).copy()
    grad = np.zeros_like(flat)
    for k in range(flat.size):
        plus = flat.copy()
        plus[k] += np.pi / 2
        minus = flat.copy()
        minus[k] -= np.pi / 2
        e_plus, _ = energy_and_grad(num_qubits, layers, field, plus,
                                    compute_gradient=False, form=form)
        e_minus, _ = energy_and_grad(num_qubits, layers, field, minus,
                                     compute_gradient=False, form=form)
        grad[k] = 0.5 * (e_plus - e_minus)
    return grad


def central_fd_grad(num_qubits: int, layers: int, field: float, params, step: float = 1e-6,
                    *, form: str = "gate") -> np.ndarray:
    flat = _flat_params(num_qubits, layers, params).copy()
    grad = np.zeros_like(flat)
    for k in range(flat.size):
        plus = flat.copy()
        plus[k] += step
        minus = flat.copy()
        minus[k] -= step
        e_plus, _ = energy_and_grad(num_qubits, layers, field, plus,
                                    compute_gradient=False, form=form)
        e_minus, _ = energy_and_grad(num_qubits, layers, field, minus,
                                     compute_gradient=False, form=form)
        grad[k] = (e_plus - e_minus) / (2 * step)
    return grad


if os.environ.get("RIO_ORACLE_THREADS"):
    set_num_threads(int(os.environ["RIO_ORACLE_THREADS"]))
