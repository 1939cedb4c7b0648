"""Plain gradient-descent loop that drives the engine.

Behaviour of ``ring_ising/training/loop.py:16-139`` in the reference: one
``step_fn`` evaluation per step, ``params <- apply_gradient_step(params, grad,
stepsize)``, per-step metrics recorded at about ``DEFAULT_PROGRESS_PARTITIONS``
evenly spaced steps when ``report_steps`` is set, and an optional one-line
progress bar.
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, Generic, TypeVar

import numpy as np

from ..config import DEFAULT_PROGRESS_PARTITIONS

ParamsT = TypeVar("ParamsT")


@dataclass(frozen=True)
class StepMetric:
    step: int
    energy: float
    grad_norm: float
    step_wall_s: float


@dataclass(frozen=True)
class StepEvaluation:
    grad: np.ndarray
    energy: float


@dataclass(frozen=True)
class LoopResult(Generic[ParamsT]):
    final_params: ParamsT
    step_metrics: tuple[StepMetric, ...]
    wall_s: float


def _progress_segment(step: int, total: int, *, partitions: int) -> int:
    """Index (1..partitions) of the progress segment that ``step`` falls in."""
    if total <= 0:
        return 0
    return min(partitions, max(1, (step * partitions - 1) // total + 1))


def _progress_line(step: int, total: int, *, elapsed_s: float, width: int = 24) -> str:
    if total <= 0:
        return "Progress [" + "-" * width + "] 0/0"
    ratio = step / total
    filled = min(width, int(round(width * ratio)))
    bar = "#" * filled + "-" * (width - filled)
    return (f"\rProgress [{bar}] {step}/{total} ({100.0 * ratio:5.1f}%) "
            f"elapsed={elapsed_s:.3f}s")


def run_gradient_descent_loop(
    params: ParamsT,
    *,
    steps: int,
    stepsize: float,
    step_fn: Callable[[ParamsT], StepEvaluation],
    apply_gradient_step: Callable[[ParamsT, np.ndarray, float], ParamsT],
    verbose: bool,
    show_progress: bool,
    report_steps: bool,
    format_step: Callable[[StepMetric], str],
) -> LoopResult[ParamsT]:
    metrics: list[StepMetric] = []
    start = time.perf_counter()
    last_segment = 0
    for step in range(1, steps + 1):
        step_start = time.perf_counter()
        segment = _progress_segment(step, steps, partitions=DEFAULT_PROGRESS_PARTITIONS)
        record = report_steps and segment != last_segment
        evaluation = step_fn(params)
        grad = np.asarray(evaluation.grad, dtype=np.float64)
        params = apply_gradient_step(params, grad, stepsize)
        wall = time.perf_counter() - step_start
        if record:
            metric = StepMetric(step=step, energy=float(evaluation.energy),
                                grad_norm=float(np.linalg.norm(grad)), step_wall_s=wall)
            metrics.append(metric)
            if verbose:
                print(format_step(metric))
        if verbose and show_progress and segment != last_segment:
            print(_progress_line(step, steps, elapsed_s=time.perf_counter() - start),
                  end="", flush=True)
        last_segment = segment
    if verbose and show_progress and steps > 0:
        print()
    return LoopResult(final_params=params, step_metrics=tuple(metrics),
                      wall_s=time.perf_counter() - start)
