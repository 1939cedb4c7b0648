"""Python entry point of the standalone engine.

Contract of ``ring_ising/backends/standalone/runtime.py:12-85``: the backend
takes a validated ``StandaloneBackendConfig``, exposes the normalised
``gradient_strategy``, ``uses_pennylane_gate_structure`` and
``estimated_workspace_gib``, and evaluates parameters of shape
(layers, num_qubits, 2) -- any other shape is a ``ValueError`` -- through
``energy``, ``energy_and_grad`` (gradient in the parameters' shape) and
``profile_energy_and_grad`` (adds the stage ``timings_ms``).  The compute runs
in ``librie.so`` on the GPU through the pybind11 module ``_cuda_backend``;
there is no CPU fallback.
"""

from __future__ import annotations

import importlib
from types import ModuleType

import numpy as np

from .config import StandaloneBackendConfig

EXTENSION_MODULE = "gpu_simulation_amd.standalone_backend._cuda_backend"
# strategies that apply the PennyLane gate sequence (RY, RZ per wire, then the
# CNOT ring) rather than fused dense layer matrices
GATE_SEQUENCE_STRATEGIES = frozenset({"inverse_walk_cuQuantum", "structured_adjoint"})


def load_extension_module() -> ModuleType:
    """Import the built pybind11 module; an unbuilt tree is an ``ImportError``."""
    try:
        return importlib.import_module(EXTENSION_MODULE)
    except ImportError as exc:
        raise ImportError("The standalone HIP extension is not built yet. "
                          "Run `python -m gpu_simulation_amd.build` first.") from exc


class RingIsingAdjointBackend:
    """One engine instance for a fixed (num_qubits, layers, field, strategy)."""

    def __init__(self, config: StandaloneBackendConfig) -> None:
        config.validate()
        self.config = config
        strategy = config.normalized_gradient_strategy
        self.gradient_strategy = strategy
        self.uses_pennylane_gate_structure = strategy in GATE_SEQUENCE_STRATEGIES
        self.estimated_workspace_gib = config.estimated_gradient_workspace_gib_for(strategy)
        ctor_args = (config.num_qubits, config.layers, float(config.field), strategy,
                     int(config.effective_structured_rotation_chunk_width),
                     bool(config.double_buffer))
        self._backend = load_extension_module()
        self._cuda = self._backend.RingIsingCudaBackend(*ctor_args)

    @property
    def native(self):
        """The underlying ``_cuda_backend.RingIsingCudaBackend`` object."""
        return self._cuda

    def _flat(self, params: np.ndarray) -> np.ndarray:
        shape = self.config.param_shape
        array = np.asarray(params, dtype=np.float64)
        if array.shape != shape:
            raise ValueError(f"Expected params with shape {shape}, got {array.shape}.")
        return np.ascontiguousarray(array).reshape(-1)

    def _evaluate(self, params: np.ndarray, *, gradient: bool, profile: bool = False) -> dict:
        return self._cuda.energy_and_grad(self._flat(params), gradient, profile)

    def _shaped(self, flat_grad) -> np.ndarray:
        return np.asarray(flat_grad, dtype=np.float64).reshape(self.config.param_shape)

    def energy(self, params: np.ndarray) -> float:
        return float(self._evaluate(params, gradient=False)["energy"])

    def energy_and_grad(self, params: np.ndarray) -> tuple[float, np.ndarray]:
        out = self._evaluate(params, gradient=True)
        return float(out["energy"]), self._shaped(out["gradient"])

    def profile_energy_and_grad(self, params: np.ndarray) -> dict[str, object]:
        """Energy, gradient and the per-stage milliseconds of one evaluation."""
        out = self._evaluate(params, gradient=True, profile=True)
        stages = out.get("timings_ms") or {}
        return {"energy": float(out["energy"]), "gradient": self._shaped(out["gradient"]),
                "timings_ms": {str(name): float(ms) for name, ms in stages.items()}}
