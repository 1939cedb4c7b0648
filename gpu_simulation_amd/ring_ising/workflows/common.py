"""Helpers shared by the workflows.

Contract of ``ring_ising/workflows/common.py:13-52``: the initial parameter
draw of a ``RunConfig``, the float64 update ``params - stepsize * grad``, and
the "Problem:" and "Summary:" blocks a verbose run prints (same headings,
labels and number formats).  The blocks are built as data -- a heading and
(label, value) rows -- and printed by one renderer, which the runtime banner
of the standalone workflow uses too.
"""

from __future__ import annotations

from typing import Iterable, Sequence

import numpy as np

from ..config import RunConfig
from ..params import make_initial_params_array
from ..training import TrainingRunResult

Rows = Sequence[tuple[str, object]]


def make_initial_params(config: RunConfig) -> np.ndarray:
    """The run's starting point, ``init_scale * N(0, 1)`` of shape (layers, num_qubits, 2)."""
    return make_initial_params_array(num_qubits=config.num_qubits, layers=config.layers,
                                     seed=config.seed, init_scale=config.init_scale)


def apply_gradient_step(params: np.ndarray, grad: np.ndarray, stepsize: float) -> np.ndarray:
    """Plain float64 gradient descent: ``params - stepsize * grad`` (a new array)."""
    out = np.array(params, dtype=np.float64)
    out -= stepsize * np.asarray(grad, dtype=np.float64)
    return out


def render_block(heading: str, rows: Iterable[tuple[str, object]]) -> str:
    """``heading:`` followed by one indented ``label: value`` line per row."""
    lines = [heading + ":"]
    lines.extend(f"  {label}: {value}" for label, value in rows)
    return "\n".join(lines)


def print_block(heading: str, rows: Iterable[tuple[str, object]], *, blank_after: bool = True) -> None:
    print(render_block(heading, rows))
    if blank_after:
        print()


def problem_rows(config: RunConfig, params: np.ndarray) -> Rows:
    return (
        ("Qubits", config.num_qubits),
        ("Layers", config.layers),
        ("Field strength", config.field),
        ("Parameter tensor shape", tuple(np.shape(params))),
    )


def result_rows(result: TrainingRunResult, config: RunConfig) -> Rows:
    rows: list[tuple[str, object]] = [("Final energy", format(result.final_energy, ".10f"))]
    if config.steps:
        rows.append(("Average step time",
                     format(1000.0 * result.wall_s / config.steps, ".3f") + " ms"))
    rows.append(("Total wall time", format(result.wall_s, ".4f") + " s"))
    return rows


def print_problem_summary(config: RunConfig, params: np.ndarray) -> None:
    print_block("Problem", problem_rows(config, params))


def print_result_summary(result: TrainingRunResult, config: RunConfig) -> None:
    print()
    print_block("Summary", result_rows(result, config), blank_after=False)
