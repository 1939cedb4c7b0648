"""Helpers shared by the workflows (``ring_ising/workflows/common.py:13-52``)."""

from __future__ import annotations

import numpy as np

from ..config import RunConfig
from ..params import make_initial_params_array
from ..training import TrainingRunResult


def make_initial_params(config: RunConfig) -> np.ndarray:
    return make_initial_params_array(num_qubits=config.num_qubits, layers=config.layers,
                                     seed=config.seed, init_scale=config.init_scale)


def apply_gradient_step(params: np.ndarray, grad: np.ndarray, stepsize: float) -> np.ndarray:
    """Plain float64 gradient descent: ``params - stepsize * grad``."""
    return np.asarray(params, dtype=np.float64) - stepsize * np.asarray(grad, dtype=np.float64)


def print_problem_summary(config: RunConfig, params: np.ndarray) -> None:
    print("Problem:")
    print(f"  Qubits: {config.num_qubits}")
    print(f"  Layers: {config.layers}")
    print(f"  Field strength: {config.field}")
    print(f"  Parameter tensor shape: {tuple(np.asarray(params).shape)}")
    print()


def print_result_summary(result: TrainingRunResult, config: RunConfig) -> None:
    print()
    print("Summary:")
    print(f"  Final energy: {result.final_energy:.10f}")
    if config.steps:
        print(f"  Average step time: {1000.0 * result.wall_s / config.steps:.3f} ms")
    print(f"  Total wall time: {result.wall_s:.4f} s")
