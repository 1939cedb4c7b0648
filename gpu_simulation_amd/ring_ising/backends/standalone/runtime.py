"""Python entry point of the standalone engine.

Mirrors ``ring_ising/backends/standalone/runtime.py:12-85`` of the reference:
params of shape (layers, n, 2) in, ``energy`` / ``energy_and_grad`` /
``profile_energy_and_grad`` out.  The compute runs in ``librie.so`` on the GPU
through the pybind11 module ``_cuda_backend``; there is no CPU fallback.
"""

from __future__ import annotations

import importlib

import numpy as np

from .config import StandaloneBackendConfig

EXTENSION_MODULE = "gpu_simulation_amd.standalone_backend._cuda_backend"


class RingIsingAdjointBackend:
    def __init__(self, config: StandaloneBackendConfig) -> None:
        config.validate()
        self.config = config
        self.gradient_strategy = config.normalized_gradient_strategy
        self.uses_pennylane_gate_structure = self.gradient_strategy in {
            "inverse_walk_cuQuantum", "structured_adjoint"}
        self.estimated_workspace_gib = config.estimated_gradient_workspace_gib_for(
            self.gradient_strategy)
        self._backend = self._load_backend()
        self._cuda = self._backend.RingIsingCudaBackend(
            config.num_qubits,
            config.layers,
            float(config.field),
            self.gradient_strategy,
            int(config.effective_structured_rotation_chunk_width),
            bool(config.double_buffer),
        )

    @staticmethod
    def _load_backend():
        try:
            return importlib.import_module(EXTENSION_MODULE)
        except ImportError as exc:
            raise ImportError(
                "The standalone HIP extension is not built yet. "
                "Run `python -m gpu_simulation_amd.build` first.") from exc

    @property
    def native(self):
        """The underlying ``_cuda_backend.RingIsingCudaBackend`` object."""
        return self._cuda

    def _normalize_params(self, params: np.ndarray) -> np.ndarray:
        array = np.asarray(params, dtype=np.float64)
        if array.shape != self.config.param_shape:
            raise ValueError(
                f"Expected params with shape {self.config.param_shape}, got {array.shape}.")
        return np.ascontiguousarray(array.reshape(-1))

    def energy(self, params: np.ndarray) -> float:
        raw = self._cuda.energy_and_grad(self._normalize_params(params), False)
        return float(raw["energy"])

    def energy_and_grad(self, params: np.ndarray) -> tuple[float, np.ndarray]:
        raw = self._cuda.energy_and_grad(self._normalize_params(params))
        grad = np.asarray(raw["gradient"], dtype=np.float64).reshape(self.config.param_shape)
        return float(raw["energy"]), grad

    def profile_energy_and_grad(self, params: np.ndarray) -> dict[str, object]:
        raw = self._cuda.energy_and_grad(self._normalize_params(params), True, True)
        timings = {str(k): float(v) for k, v in dict(raw.get("timings_ms", {})).items()}
        return {
            "energy": float(raw["energy"]),
            "gradient": np.asarray(raw["gradient"], dtype=np.float64).reshape(
                self.config.param_shape),
            "timings_ms": timings,
        }
