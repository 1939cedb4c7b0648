"""CPU tests of the gradient-descent loop and the workflow's printed blocks.

The loop's contract (``ring_ising/training/loop.py:75-139`` in the reference):
one ``step_fn`` call and one update per step; the energy recorded for a step
is the one evaluated before its update; with ``report_steps`` a metric is kept
for step s exactly when ceil(s * P / T) > ceil((s - 1) * P / T), P = 60
partitions, T = steps -- the first step of each progress segment.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from gpu_simulation_amd.ring_ising.config import DEFAULT_PROGRESS_PARTITIONS, RunConfig
from gpu_simulation_amd.ring_ising.training import (
    StepEvaluation,
    TrainingRunResult,
    run_gradient_descent_loop,
)
from gpu_simulation_amd.ring_ising.training.loop import _ProgressBar, marked_steps
from gpu_simulation_amd.ring_ising.workflows import common
from gpu_simulation_amd.ring_ising.workflows.standalone import _backend_config


def _segment_starts(total: int, partitions: int) -> set[int]:
    def seg(s: int) -> int:
        return 0 if s == 0 else math.ceil(s * partitions / total)
    return {s for s in range(1, total + 1) if seg(s) > seg(s - 1)}


@pytest.mark.parametrize("total", [0, 1, 2, 7, 20, 59, 60, 61, 119, 120, 121, 200, 1000, 12345])
def test_marked_steps_are_segment_starts(total):
    assert marked_steps(total, DEFAULT_PROGRESS_PARTITIONS) == _segment_starts(total, 60)


def test_marked_steps_small_cases():
    assert marked_steps(20, 60) == set(range(1, 21))
    assert marked_steps(120, 60) == set(range(1, 120, 2))
    assert len(marked_steps(12345, 60)) == 60


def _quadratic_loop(steps, report, stepsize=0.25):
    """Descent on E(p) = |p|^2 / 2, grad = p: p_k = (1 - stepsize)^k p_0."""
    calls = []

    def step_fn(p):
        calls.append(p.copy())
        return StepEvaluation(grad=p.copy(), energy=0.5 * float(p @ p))

    p0 = np.array([1.0, -2.0, 0.5])
    res = run_gradient_descent_loop(
        p0, steps=steps, stepsize=stepsize, step_fn=step_fn,
        apply_gradient_step=common.apply_gradient_step, verbose=False, show_progress=False,
        report_steps=report, format_step=str)
    return p0, calls, res


def test_loop_updates_and_records_pre_update_energy():
    p0, calls, res = _quadratic_loop(130, True)
    assert len(calls) == 130
    np.testing.assert_allclose(res.final_params, 0.75 ** 130 * p0, rtol=1e-12, atol=0)
    assert [m.step for m in res.step_metrics] == sorted(_segment_starts(130, 60))
    for m in res.step_metrics:
        p_before = 0.75 ** (m.step - 1) * p0
        assert m.energy == pytest.approx(0.5 * p_before @ p_before, rel=1e-12)
        assert m.grad_norm == pytest.approx(np.linalg.norm(p_before), rel=1e-12)
        assert m.step_wall_s >= 0.0
    assert res.wall_s >= 0.0


def test_loop_without_reporting_keeps_no_metrics():
    _, calls, res = _quadratic_loop(5, False)
    assert len(calls) == 5 and res.step_metrics == ()


def test_zero_steps_returns_the_start():
    p0, calls, res = _quadratic_loop(0, True)
    assert calls == [] and res.step_metrics == ()
    np.testing.assert_array_equal(res.final_params, p0)


def test_progress_line_format():
    bar = _ProgressBar(8)
    line = bar.line(3)
    assert line.startswith("\rProgress [#########---------------] 3/8 ( 37.5%) elapsed=")
    assert line.endswith("s")
    assert _ProgressBar(0).line(0) == "Progress [------------------------] 0/0"
    assert _ProgressBar(4).line(4).startswith("\rProgress [" + "#" * 24 + "] 4/4 (100.0%)")


def test_verbose_loop_prints_recorded_steps(capsys):
    def step_fn(p):
        return StepEvaluation(grad=np.ones(2), energy=1.5)

    run_gradient_descent_loop(np.zeros(2), steps=3, stepsize=0.1, step_fn=step_fn,
                              apply_gradient_step=common.apply_gradient_step, verbose=True,
                              show_progress=False, report_steps=True,
                              format_step=lambda m: f"step={m.step}")
    assert capsys.readouterr().out.splitlines() == ["step=1", "step=2", "step=3"]


def test_apply_gradient_step_is_float64_and_new():
    p = np.array([1.0, 2.0], dtype=np.float32)
    out = common.apply_gradient_step(p, np.array([1, 1]), 0.5)
    assert out.dtype == np.float64 and out is not p
    np.testing.assert_array_equal(out, [0.5, 1.5])
    np.testing.assert_array_equal(p, [1.0, 2.0])


def test_problem_and_summary_blocks(capsys):
    cfg = RunConfig(num_qubits=6, layers=2, field=0.5, steps=4)
    common.print_problem_summary(cfg, np.zeros((2, 6, 2)))
    res = TrainingRunResult(backend_label="standalone_hip", final_params=np.zeros(1),
                            final_energy=-1.25, step_metrics=(), wall_s=0.5)
    common.print_result_summary(res, cfg)
    assert capsys.readouterr().out.splitlines() == [
        "Problem:", "  Qubits: 6", "  Layers: 2", "  Field strength: 0.5",
        "  Parameter tensor shape: (2, 6, 2)", "",
        "", "Summary:", "  Final energy: -1.2500000000", "  Average step time: 125.000 ms",
        "  Total wall time: 0.5000 s"]


def test_summary_without_steps_omits_the_average(capsys):
    res = TrainingRunResult(backend_label="x", final_params=np.zeros(1), final_energy=0.0,
                            step_metrics=(), wall_s=0.0)
    common.print_result_summary(res, RunConfig(steps=0))
    assert "Average step time" not in capsys.readouterr().out


def test_backend_config_follows_run_config():
    cfg = RunConfig(num_qubits=9, layers=4, field=0.3, gradient_strategy="inverse_walk_cuquantum",
                    structured_rotation_chunk_width=3, mode2_rotation_chunk_width=5,
                    double_buffer=True)
    bc = _backend_config(cfg)
    assert (bc.num_qubits, bc.layers, bc.field) == (9, 4, 0.3)
    assert bc.gradient_strategy == "inverse_walk_cuquantum"
    assert bc.structured_rotation_chunk_width == 5  # the legacy field wins when set
    assert bc.double_buffer is True
