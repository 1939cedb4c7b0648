"""Plain gradient-descent loop that drives the engine.

Contract of ``ring_ising/training/loop.py:16-139`` in the reference: the
``StepMetric`` / ``StepEvaluation`` / ``LoopResult`` records, one ``step_fn``
evaluation and one ``apply_gradient_step`` update per step, and -- when
``report_steps`` is set -- a metric for the first step of each of
``DEFAULT_PROGRESS_PARTITIONS`` equal shares of the run (every step when the
run is shorter than that).  The optional progress bar is redrawn on the same
steps.

Here the recorded steps are computed once before the loop (``marked_steps``),
and the progress bar is a small object that owns its clock and its line.
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
    """Energy, gradient norm and wall time of one recorded step."""

    step: int
    energy: float
    grad_norm: float
    step_wall_s: float


@dataclass(frozen=True)
class StepEvaluation:
    """What ``step_fn`` returns: the gradient and the energy at the same parameters."""

    grad: np.ndarray
    energy: float


@dataclass(frozen=True)
class LoopResult(Generic[ParamsT]):
    """Parameters after the last update, the recorded metrics and the loop's wall time."""

    final_params: ParamsT
    step_metrics: tuple[StepMetric, ...]
    wall_s: float


def marked_steps(total: int, partitions: int) -> frozenset[int]:
    """Steps (1-based) that open one of ``partitions`` equal shares of ``total`` steps.

    Share k (k = 1..partitions) covers the steps s with (k-1)/partitions <
    s/total <= k/partitions; its first step is floor((k-1) * total / partitions)
    + 1.  Shares that hold no step (total < partitions) repeat a step, so the
    set simply holds every step then.
    """
    if total <= 0 or partitions <= 0:
        return frozenset()
    return frozenset((k * total) // partitions + 1 for k in range(partitions))


class _ProgressBar:
    """One-line ``Progress [###---] i/n (pct%) elapsed=...s`` bar, redrawn in place."""

    WIDTH = 24

    def __init__(self, total: int) -> None:
        self.total = total
        self.started = time.perf_counter()

    def line(self, done: int) -> str:
        if self.total <= 0:
            return "Progress [" + "-" * self.WIDTH + "] 0/0"
        frac = done / self.total
        hashes = min(self.WIDTH, int(round(self.WIDTH * frac)))
        bar = "".join("#" if i < hashes else "-" for i in range(self.WIDTH))
        pct = format(100.0 * frac, "5.1f")
        elapsed = format(time.perf_counter() - self.started, ".3f")
        return "\rProgress [%s] %d/%d (%s%%) elapsed=%ss" % (bar, done, self.total, pct, elapsed)

    def draw(self, done: int) -> None:
        print(self.line(done), end="", flush=True)


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
    """Run ``steps`` plain gradient-descent updates of ``params``.

    The update is whatever ``apply_gradient_step(params, grad, stepsize)``
    returns; the energy recorded for a step is the one ``step_fn`` reported at
    the parameters *before* that step's update.
    """
    marks = marked_steps(steps, DEFAULT_PROGRESS_PARTITIONS)
    bar = _ProgressBar(steps) if (verbose and show_progress) else None
    recorded: list[StepMetric] = []
    t_loop = bar.started if bar is not None else time.perf_counter()

    for step in range(1, steps + 1):
        t_step = time.perf_counter()
        ev = step_fn(params)
        grad = np.asarray(ev.grad, dtype=np.float64)
        params = apply_gradient_step(params, grad, stepsize)
        took = time.perf_counter() - t_step
        if step not in marks:
            continue
        if report_steps:
            metric = StepMetric(step, float(ev.energy), float(np.linalg.norm(grad)), took)
            recorded.append(metric)
            if verbose:
                print(format_step(metric))
        if bar is not None:
            bar.draw(step)

    if bar is not None and steps > 0:
        print()
    return LoopResult(params, tuple(recorded), time.perf_counter() - t_loop)
