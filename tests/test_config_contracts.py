"""Configuration contracts of the Python surface (CPU only).

Mirrors the reference's CPU-side tests (tests/test_backends_parity.py:237-292,
tests/test_training_workflow_interfaces.py:18-20): removed strategy names and
out-of-range widths raise ValueError, dense_scan is limited to 8 qubits, and
the dense workspace estimate follows the layer-level operator count.
"""

from __future__ import annotations

import pytest

from gpu_simulation_amd.ring_ising import BACKENDS, RunConfig, StandaloneBackendConfig
from gpu_simulation_amd.ring_ising.config import (
    STANDALONE_GRADIENT_STRATEGIES,
    SUPPORTED_STANDALONE_GRADIENT_STRATEGIES,
)
from gpu_simulation_amd.ring_ising.training import StepMetric

REMOVED_MODES = ("inverse_walk", "ryrz_fused", "mode2", "save_param_states", "checkpoint",
                 "block_fused_adjoint", "intrablock_parallel", "mode2_chunk2", "mode2_chunk3",
                 "mode2_dense2", "mode2_dense3")


def test_backend_names_and_strategies():
    assert BACKENDS == ("pennylane", "standalone")
    assert STANDALONE_GRADIENT_STRATEGIES == ("inverse_walk_cuQuantum", "structured_adjoint",
                                              "dense_scan")
    assert "inverse_walk_cuquantum" in SUPPORTED_STANDALONE_GRADIENT_STRATEGIES
    assert StepMetric.__name__ == "StepMetric"


def test_run_config_defaults():
    cfg = RunConfig()
    assert (cfg.backend, cfg.num_qubits, cfg.layers, cfg.field, cfg.steps, cfg.stepsize,
            cfg.seed, cfg.init_scale) == ("standalone", 12, 3, 1.0, 20, 0.08, 7, 0.3)
    assert cfg.gradient_strategy == "structured_adjoint"
    assert cfg.effective_structured_rotation_chunk_width == 8
    assert RunConfig(mode2_rotation_chunk_width=3).effective_structured_rotation_chunk_width == 3


@pytest.mark.parametrize("mode", REMOVED_MODES)
def test_removed_legacy_modes_are_rejected(mode):
    with pytest.raises(ValueError):
        StandaloneBackendConfig(num_qubits=4, layers=2, field=1.0, gradient_strategy=mode).validate()


def test_alias_is_normalised():
    cfg = StandaloneBackendConfig(num_qubits=4, layers=2, field=1.0,
                                  gradient_strategy="inverse_walk_cuquantum")
    cfg.validate()
    assert cfg.normalized_gradient_strategy == "inverse_walk_cuQuantum"


def test_dense_scan_rejects_more_than_eight_qubits():
    with pytest.raises(ValueError):
        StandaloneBackendConfig(num_qubits=9, layers=2, field=1.0,
                                gradient_strategy="dense_scan").validate()
    StandaloneBackendConfig(num_qubits=8, layers=2, field=1.0, gradient_strategy="dense_scan").validate()


def test_dense_scan_workspace_estimate_uses_layer_level_ops():
    cfg = StandaloneBackendConfig(num_qubits=6, layers=32, field=1.0, gradient_strategy="dense_scan")
    num_ops = cfg.layers * 2
    padded = 1 << (num_ops - 1).bit_length()
    matrix_count = num_ops + padded + max(0, padded - 1) + 2
    vector_count = 3 * padded + max(num_ops + 1, padded) + 3
    expected = (matrix_count * cfg.dense_matrix_nbytes + vector_count * cfg.statevector_nbytes
                + cfg.num_params * (8 + 8 + 4)) / (1024 ** 3)
    assert cfg.estimated_gradient_state_buffers_for("dense_scan") == vector_count
    assert cfg.estimated_gradient_workspace_gib_for("dense_scan") == pytest.approx(expected)
    assert cfg.estimated_gradient_state_buffers_for("structured_adjoint") == 4


@pytest.mark.parametrize("width", [0, 9])
def test_structured_rotation_chunk_width_is_validated(width):
    with pytest.raises(ValueError):
        StandaloneBackendConfig(num_qubits=4, layers=2, field=1.0,
                                gradient_strategy="structured_adjoint",
                                structured_rotation_chunk_width=width).validate()


@pytest.mark.parametrize("kwargs", [dict(num_qubits=1, layers=2), dict(num_qubits=4, layers=0)])
def test_problem_size_is_validated(kwargs):
    with pytest.raises(ValueError):
        StandaloneBackendConfig(field=1.0, gradient_strategy="structured_adjoint", **kwargs).validate()


def test_pennylane_backend_is_not_silently_substituted():
    from gpu_simulation_amd.ring_ising import run

    with pytest.raises(NotImplementedError):
        run(RunConfig(backend="pennylane", verbose=False))
    with pytest.raises(ValueError):
        run(RunConfig(backend="cuda", verbose=False))


@pytest.mark.parametrize("kwargs", [
    {"drift_tol": 1e-12, "schedule": "stored"},
    {"drift_tol": 1e-12, "gradient_strategy": "dense_scan"},
    {"drift_tol": -1.0},
    {"drift_tol": float("nan")},
    {"drift_tol": 1e-12, "drift_check_every": -1},
])
def test_drift_triggered_checkpoints_reject_bad_settings(kwargs):
    """Checked before any engine is created, so no GPU is needed."""
    from gpu_simulation_amd.autograd import RingIsingEnergy

    with pytest.raises(ValueError):
        RingIsingEnergy(6, 2, 1.0, **kwargs)
