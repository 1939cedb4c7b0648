"""Python surface of the ring-Ising workflow, backed by the MI355X engine.

Names follow the reference package ``ring_ising`` (``ring_ising/__init__.py``):
``RunConfig``, ``run``, ``run_standalone``, ``RingIsingAdjointBackend``,
``StandaloneBackendConfig``, ``StepMetric``, ``TrainingRunResult`` ...
"""

from __future__ import annotations

from .backends.standalone import RingIsingAdjointBackend, StandaloneBackendConfig
from .config import (
    BACKENDS,
    DEFAULT_PROGRESS_PARTITIONS,
    STANDALONE_GRADIENT_STRATEGIES,
    SUPPORTED_STANDALONE_GRADIENT_STRATEGIES,
    RunConfig,
)
from .training import StepEvaluation, StepMetric, TrainingRunResult
from .workflows import run, run_pennylane, run_standalone
from .workflows.standalone import StandaloneResult, StandaloneWorkflow

__all__ = [
    "BACKENDS",
    "DEFAULT_PROGRESS_PARTITIONS",
    "RunConfig",
    "STANDALONE_GRADIENT_STRATEGIES",
    "SUPPORTED_STANDALONE_GRADIENT_STRATEGIES",
    "RingIsingAdjointBackend",
    "StandaloneBackendConfig",
    "StandaloneResult",
    "StandaloneWorkflow",
    "StepEvaluation",
    "StepMetric",
    "TrainingRunResult",
    "run",
    "run_pennylane",
    "run_standalone",
]
