"""Backend configuration and workspace estimates.

Mirrors ``ring_ising/backends/standalone/config.py:17-117`` of the reference:
the same validation order and messages (ValueError), and the same workspace
formulas, which callers use for sizing decisions.
"""

from __future__ import annotations

from dataclasses import dataclass

from ...config import (
    STANDALONE_GRADIENT_STRATEGIES,
    STANDALONE_GRADIENT_STRATEGY_ALIASES,
    STRUCTURED_ROTATION_CHUNK_WIDTH_MAX,
    SUPPORTED_STANDALONE_GRADIENT_STRATEGIES,
)

STRUCTURED_ADJOINT_GRADIENT_STRATEGIES = {"structured_adjoint"}


def _padded_op_count(layers: int) -> tuple[int, int]:
    """(number of dense operators, next power of two) for 2 ops per layer."""
    num_ops = 2 * layers
    padded = 1 << (num_ops - 1).bit_length() if num_ops > 0 else 1
    return num_ops, padded


@dataclass(frozen=True)
class StandaloneBackendConfig:
    num_qubits: int
    layers: int
    field: float
    gradient_strategy: str
    structured_rotation_chunk_width: int = 8
    mode2_rotation_chunk_width: int | None = None
    double_buffer: bool = False

    @property
    def param_shape(self) -> tuple[int, int, int]:
        return (self.layers, self.num_qubits, 2)

    @property
    def num_params(self) -> int:
        return 2 * self.layers * self.num_qubits

    @property
    def normalized_gradient_strategy(self) -> str:
        name = self.gradient_strategy
        return STANDALONE_GRADIENT_STRATEGY_ALIASES.get(name, name)

    @property
    def effective_structured_rotation_chunk_width(self) -> int:
        if self.mode2_rotation_chunk_width is not None:
            return self.mode2_rotation_chunk_width
        return self.structured_rotation_chunk_width

    @property
    def statevector_nbytes(self) -> int:
        return 16 << self.num_qubits

    @property
    def dense_matrix_nbytes(self) -> int:
        dim = 1 << self.num_qubits
        return 16 * dim * dim

    def estimated_gradient_state_buffers_for(self, strategy: str) -> int:
        name = STANDALONE_GRADIENT_STRATEGY_ALIASES.get(strategy, strategy)
        if name == "inverse_walk_cuQuantum" or name in STRUCTURED_ADJOINT_GRADIENT_STRATEGIES:
            return 4
        if name == "dense_scan":
            num_ops, padded = _padded_op_count(self.layers)
            return 3 * padded + max(num_ops + 1, padded) + 3
        raise ValueError("strategy must be one of STANDALONE_GRADIENT_STRATEGIES.")

    def estimated_gradient_workspace_gib_for(self, strategy: str) -> float:
        name = STANDALONE_GRADIENT_STRATEGY_ALIASES.get(strategy, strategy)
        if name == "dense_scan":
            num_ops, padded = _padded_op_count(self.layers)
            if num_ops <= 0:
                return 0.0
            matrices = num_ops + padded + max(0, padded - 1) + 2
            vectors = 3 * padded + max(num_ops + 1, padded) + 3
            total = (matrices * self.dense_matrix_nbytes + vectors * self.statevector_nbytes
                     + self.num_params * (8 + 8 + 4))
            return total / float(1 << 30)
        buffers = self.estimated_gradient_state_buffers_for(strategy)
        return buffers * self.statevector_nbytes / float(1 << 30)

    def validate(self) -> None:
        if self.num_qubits < 2:
            raise ValueError("num_qubits must be at least 2.")
        if self.layers < 1:
            raise ValueError("layers must be at least 1.")
        if self.gradient_strategy not in SUPPORTED_STANDALONE_GRADIENT_STRATEGIES:
            raise ValueError(
                "gradient_strategy must be one of public strategies "
                f"{STANDALONE_GRADIENT_STRATEGIES!r}; experimental/legacy accepted strategies "
                f"are {SUPPORTED_STANDALONE_GRADIENT_STRATEGIES!r}.")
        if self.normalized_gradient_strategy == "dense_scan" and self.num_qubits > 8:
            raise ValueError("dense_scan requires num_qubits <= 8.")
        width = self.effective_structured_rotation_chunk_width
        if width < 1:
            raise ValueError("structured_rotation_chunk_width must be at least 1.")
        if width > STRUCTURED_ROTATION_CHUNK_WIDTH_MAX:
            raise ValueError("structured_rotation_chunk_width must be at most "
                             f"{STRUCTURED_ROTATION_CHUNK_WIDTH_MAX}.")
