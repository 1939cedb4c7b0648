"""Run results and shared argument checks (``ring_ising/training/result.py``).

GPU telemetry in the reference polls ``nvidia-smi``; on this platform kernel
time and HBM traffic come from ``rocprofv3`` instead, so the telemetry hook
runs the body and returns no summary.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from typing import Any, Callable, TypeVar

import numpy as np

from .loop import StepMetric

ResultT = TypeVar("ResultT")


@dataclass
class TrainingRunResult:
    backend_label: str
    final_params: np.ndarray
    final_energy: float
    step_metrics: tuple[StepMetric, ...]
    wall_s: float
    gpu_telemetry_summary: Any | None = None
    metadata: dict[str, Any] = field(default_factory=dict)


def validate_common_run_args(*, num_qubits: int, layers: int, steps: int,
                             telemetry_interval_s: float) -> None:
    if num_qubits < 2:
        raise ValueError("num_qubits must be at least 2 for the ring Ising Hamiltonian.")
    if layers < 1:
        raise ValueError("layers must be at least 1.")
    if steps < 0:
        raise ValueError("steps must be non-negative.")
    if telemetry_interval_s <= 0:
        raise ValueError("telemetry_interval_s must be positive.")


def run_with_optional_telemetry(*, enabled: bool, interval_s: float, live: bool, label: str,
                                body: Callable[[], ResultT]) -> tuple[ResultT, None]:
    del interval_s, live
    if enabled:
        warnings.warn(f"{label}: nvidia-smi telemetry is not available on AMD GPUs; "
                      "profile with rocprofv3 instead.", RuntimeWarning, stacklevel=2)
    return body(), None
