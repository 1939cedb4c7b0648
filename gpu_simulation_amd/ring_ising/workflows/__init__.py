"""Workflow dispatch (``ring_ising/workflows/__init__.py:11-19``).

The ``"pennylane"`` backend name is accepted by the configuration for
compatibility, but PennyLane's NVIDIA ``lightning.gpu`` device is not part of
this build: selecting it raises ``NotImplementedError``.
"""

from __future__ import annotations

from ..config import BACKENDS, RunConfig
from .standalone import run_standalone


def run_pennylane(config: RunConfig):
    raise NotImplementedError(
        "The PennyLane lightning.gpu workflow is not available on this platform; "
        "use backend='standalone'.")


def run(config: RunConfig):
    if config.backend == "pennylane":
        return run_pennylane(config)
    if config.backend == "standalone":
        return run_standalone(config)
    raise ValueError(f"Unknown backend {config.backend!r}. Expected one of: {', '.join(BACKENDS)}")


__all__ = ["BACKENDS", "RunConfig", "run", "run_pennylane", "run_standalone"]
