"""Unified workflow CLI with the reference's flags (``ring_ising/cli/run.py:24-108``,
``ring_ising/cli/common.py:10-110``).

``python -m gpu_simulation_amd.ring_ising.cli --backend standalone --qubits 20 --layers 2``
runs the standalone gradient-descent workflow on the MI355X engine.  Every flag of the
reference is accepted with the same name, default and meaning; ``--backend pennylane``
raises (PennyLane's lightning.gpu is NVIDIA-only) and the GPU telemetry flags warn, as
the rest of this package does.
"""

from __future__ import annotations

import argparse
from typing import Sequence

from ..config import (
    BACKENDS,
    STANDALONE_GRADIENT_STRATEGIES,
    STRUCTURED_ROTATION_CHUNK_WIDTH_MAX,
    SUPPORTED_STANDALONE_GRADIENT_STRATEGIES,
    RunConfig,
)
from ..workflows import run

DEFAULT_RUN_CONFIG = RunConfig()


def parse_args(argv: Sequence[str] | None = None) -> argparse.Namespace:
    parser = argparse.ArgumentParser(
        description=("Run the ring-Ising experiment with the standalone backend on an AMD "
                     "Instinct MI355X (the PennyLane backend needs lightning.gpu)."))
    parser.add_argument("--backend", choices=BACKENDS, default="standalone",
                        help="Backend to execute: 'pennylane' or 'standalone'.")
    parser.add_argument("--qubits", dest="num_qubits", type=int,
                        default=DEFAULT_RUN_CONFIG.num_qubits, help="Number of qubits.")
    parser.add_argument("--layers", type=int, default=DEFAULT_RUN_CONFIG.layers,
                        help="Ansatz layers.")
    parser.add_argument("--field", type=float, default=DEFAULT_RUN_CONFIG.field,
                        help="Transverse-field Ising X-field strength.")
    parser.add_argument("--steps", type=int, default=DEFAULT_RUN_CONFIG.steps,
                        help="Measured gradient steps.")
    parser.add_argument("--stepsize", type=float, default=DEFAULT_RUN_CONFIG.stepsize,
                        help="Manual gradient-descent step size.")
    parser.add_argument("--seed", type=int, default=DEFAULT_RUN_CONFIG.seed,
                        help="Random seed for parameter initialization.")
    parser.add_argument("--init-scale", dest="init_scale", type=float,
                        default=DEFAULT_RUN_CONFIG.init_scale,
                        help="Standard deviation for the initial parameter distribution.")
    parser.add_argument("--gradient-strategy", choices=SUPPORTED_STANDALONE_GRADIENT_STRATEGIES,
                        default=DEFAULT_RUN_CONFIG.gradient_strategy,
                        help=("Standalone adjoint strategy. Public strategies: "
                              f"{', '.join(STANDALONE_GRADIENT_STRATEGIES)}."))
    parser.add_argument("--structured-rotation-chunk-width", type=int,
                        choices=range(1, STRUCTURED_ROTATION_CHUNK_WIDTH_MAX + 1),
                        default=DEFAULT_RUN_CONFIG.structured_rotation_chunk_width,
                        metavar=f"1..{STRUCTURED_ROTATION_CHUNK_WIDTH_MAX}",
                        help="Rotation-layer fusion width for structured_adjoint.")
    parser.add_argument("--mode2-rotation-chunk-width", dest="structured_rotation_chunk_width",
                        type=int, choices=range(1, STRUCTURED_ROTATION_CHUNK_WIDTH_MAX + 1),
                        help=argparse.SUPPRESS)
    parser.add_argument("--double-buffer", dest="double_buffer", action="store_true",
                        default=DEFAULT_RUN_CONFIG.double_buffer,
                        help="Accepted for compatibility (structured_adjoint only).")
    parser.add_argument("--report-steps", action="store_true",
                        default=DEFAULT_RUN_CONFIG.report_steps,
                        help="Print step reports at about 60 evenly spaced checkpoints.")
    parser.add_argument("--gpu-telemetry", action="store_true",
                        default=DEFAULT_RUN_CONFIG.gpu_telemetry,
                        help="Accepted for compatibility; use rocprofv3 on AMD GPUs.")
    parser.add_argument("--telemetry-interval", dest="telemetry_interval_s", type=float,
                        default=DEFAULT_RUN_CONFIG.telemetry_interval_s,
                        help="Sampling interval in seconds for GPU telemetry.")
    parser.add_argument("--telemetry-live", action="store_true",
                        default=DEFAULT_RUN_CONFIG.telemetry_live,
                        help="Print live telemetry lines (accepted for compatibility).")
    return parser.parse_args(argv)


def main(argv: Sequence[str] | None = None):
    args = parse_args(argv)
    return run(RunConfig(**vars(args)))


if __name__ == "__main__":
    main()
