"""CLI flag surface and the stage-profiling harness (CPU parts).

Mirrors the reference's ``ring_ising/cli/run.py`` flags and
``benchmarks/profile_baseline_stages.py`` case parsing / CSV columns.
"""

from __future__ import annotations

import argparse

import pytest

from gpu_simulation_amd import profiling
from gpu_simulation_amd.ring_ising.cli import parse_args
from gpu_simulation_amd.ring_ising.config import RunConfig


def test_cli_defaults_are_run_config_defaults():
    args = vars(parse_args([]))
    cfg = RunConfig(**args)
    default = RunConfig()
    for key in ("backend", "num_qubits", "layers", "field", "steps", "stepsize", "seed",
                "init_scale", "gradient_strategy", "structured_rotation_chunk_width",
                "double_buffer", "report_steps", "gpu_telemetry", "telemetry_interval_s",
                "telemetry_live"):
        assert getattr(cfg, key) == getattr(default, key), key


def test_cli_flags_and_aliases():
    a = parse_args(["--qubits", "16", "--layers", "4", "--field", "0.5", "--steps", "3",
                    "--stepsize", "0.1", "--seed", "1", "--init-scale", "0.2",
                    "--gradient-strategy", "inverse_walk_cuquantum",
                    "--mode2-rotation-chunk-width", "4", "--double-buffer"])
    assert (a.num_qubits, a.layers, a.field, a.steps, a.stepsize, a.seed, a.init_scale) == (
        16, 4, 0.5, 3, 0.1, 1, 0.2)
    assert a.gradient_strategy == "inverse_walk_cuquantum"
    assert a.structured_rotation_chunk_width == 4
    assert a.double_buffer is True
    for bad in (["--gradient-strategy", "ryrz_fused"], ["--structured-rotation-chunk-width", "9"],
                ["--backend", "cuda"]):
        with pytest.raises(SystemExit):
            parse_args(bad)


def test_cli_pennylane_backend_is_unavailable():
    from gpu_simulation_amd.ring_ising.cli import main

    with pytest.raises(NotImplementedError):
        main(["--backend", "pennylane", "--qubits", "4", "--layers", "1", "--steps", "0"])


def test_profiling_case_parser_and_columns():
    assert profiling.parse_case("12x16") == profiling.BenchmarkCase(12, 16)
    with pytest.raises(argparse.ArgumentTypeError):
        profiling.parse_case("12-16")
    args = profiling.parse_args([])
    assert [(c.num_qubits, c.layers) for c in args.cases][:3] == [(12, 8), (12, 32), (12, 128)]
    assert args.structured_widths == [1, 2, 3, 4, 8]
    assert profiling.CSV_FIELDS == ["num_qubits", "layers", "mode", "gradient_strategy",
                                    "structured_rotation_chunk_width", "stage", "median_ms",
                                    "total_stage_ms"]
