"""Initial parameters: ``init_scale * N(0, 1)`` of shape (layers, n, 2) from
``numpy.random.default_rng(seed)`` (reference: ``ring_ising/params.py:8-20``)."""

from __future__ import annotations

import numpy as np


def make_initial_params_array(*, num_qubits: int, layers: int, seed: int,
                              init_scale: float) -> np.ndarray:
    rng = np.random.default_rng(seed)
    draws = rng.standard_normal((layers, num_qubits, 2))
    return np.asarray(init_scale * draws, dtype=np.float64)
