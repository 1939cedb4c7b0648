"""Command-line entry point of the ring-Ising workflow (``python -m gpu_simulation_amd.ring_ising.cli``)."""

from .run import main, parse_args

__all__ = ["main", "parse_args"]
