"""Standalone (HIP) training workflow.

Contract of ``ring_ising/workflows/standalone.py:25-167``: ``run_standalone``
accepts only ``backend="standalone"``, checks the shared run arguments, builds
the engine from the ``RunConfig``, draws the initial parameters, runs the
plain gradient-descent loop with one ``energy_and_grad`` per step, evaluates
the energy of the final parameters and returns a ``TrainingRunResult`` whose
metadata names the gradient strategy, the chunk width, the workspace estimate
and the gate structure.  A verbose run prints the runtime banner, the problem
block, the initial energy, one ``step=...`` line per recorded step and the
summary block.
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from functools import partial

import numpy as np

from ..backends.standalone import RingIsingAdjointBackend, StandaloneBackendConfig
from ..config import RunConfig
from ..training import (
    StepEvaluation,
    StepMetric,
    TrainingRunResult,
    run_gradient_descent_loop,
    run_with_optional_telemetry,
    validate_common_run_args,
)
from .common import (
    apply_gradient_step,
    make_initial_params,
    print_block,
    print_problem_summary,
    print_result_summary,
)

BACKEND_LABEL = "standalone_hip"
# RunConfig fields handed to the engine's configuration unchanged
_ENGINE_FIELDS = ("num_qubits", "layers", "field", "gradient_strategy", "double_buffer")


@dataclass(frozen=True)
class StandaloneWorkflow:
    """A run's configuration, the engine configuration derived from it and the engine."""

    config: RunConfig
    backend_config: StandaloneBackendConfig
    backend: RingIsingAdjointBackend

    @property
    def chunk_width(self) -> int:
        return self.backend_config.effective_structured_rotation_chunk_width

    def evaluate(self, params: np.ndarray) -> StepEvaluation:
        energy, grad = self.backend.energy_and_grad(params)
        return StepEvaluation(grad=np.asarray(grad, dtype=np.float64), energy=float(energy))

    def metadata(self) -> dict[str, object]:
        eng = self.backend
        return {
            "gradient_strategy": eng.gradient_strategy,
            "structured_rotation_chunk_width": self.chunk_width,
            "estimated_workspace_gib": eng.estimated_workspace_gib,
            "pennylane_gate_structure": eng.uses_pennylane_gate_structure,
        }

    def runtime_rows(self) -> list[tuple[str, object]]:
        eng = self.backend
        rows: list[tuple[str, object]] = [("Backend", "Standalone HIP (MI355X)"),
                                          ("Gradient strategy", eng.gradient_strategy)]
        if eng.gradient_strategy == "structured_adjoint":
            rows.append(("Structured rotation chunk width", self.chunk_width))
        structure = ("PennyLane-style gate structure" if eng.uses_pennylane_gate_structure
                     else "Dense-scan fused gate structure")
        rows.append((structure, True))
        rows.append(("Estimated gradient workspace", f"{eng.estimated_workspace_gib:.2f} GiB"))
        return rows


StandaloneResult = TrainingRunResult


def _backend_config(config: RunConfig) -> StandaloneBackendConfig:
    shared = {name: getattr(config, name) for name in _ENGINE_FIELDS}
    return StandaloneBackendConfig(
        structured_rotation_chunk_width=config.effective_structured_rotation_chunk_width,
        **shared)


def create_standalone_workflow(config: RunConfig) -> StandaloneWorkflow:
    backend_config = _backend_config(config)
    return StandaloneWorkflow(config, backend_config, RingIsingAdjointBackend(backend_config))


def print_standalone_runtime_summary(workflow: StandaloneWorkflow, params: np.ndarray) -> None:
    print_block("Runtime", workflow.runtime_rows())
    print_problem_summary(workflow.config, params)
    details: list[tuple[str, object]] = [
        ("Initial energy", f"{workflow.backend.energy(params):.10f}")]
    if workflow.config.gpu_telemetry:
        details.append(("GPU telemetry", "not sampled on this platform (use rocprofv3)"))
    print_block("Backend details", details)


def _format_step(metric: StepMetric) -> str:
    return "step=%03d energy=%.10f grad_norm=%.10f step_ms=%.3f" % (
        metric.step, metric.energy, metric.grad_norm, 1000.0 * metric.step_wall_s)


def _train(workflow: StandaloneWorkflow, start: np.ndarray) -> TrainingRunResult:
    cfg = workflow.config
    t0 = time.perf_counter()
    loop = run_gradient_descent_loop(
        start,
        steps=cfg.steps,
        stepsize=cfg.stepsize,
        step_fn=workflow.evaluate,
        apply_gradient_step=apply_gradient_step,
        verbose=cfg.verbose,
        show_progress=cfg.show_progress and not cfg.report_steps,
        report_steps=cfg.report_steps,
        format_step=_format_step,
    )
    final = np.asarray(loop.final_params, dtype=np.float64)
    final_energy = float(workflow.backend.energy(final))
    return TrainingRunResult(backend_label=BACKEND_LABEL, final_params=final,
                             final_energy=final_energy, step_metrics=loop.step_metrics,
                             wall_s=time.perf_counter() - t0, metadata=workflow.metadata())


def run_standalone(config: RunConfig) -> TrainingRunResult:
    """Run one standalone training workflow end to end (see the module docstring)."""
    if config.backend != "standalone":
        raise ValueError(f"run_standalone expected backend='standalone', got {config.backend!r}.")
    validate_common_run_args(num_qubits=config.num_qubits, layers=config.layers,
                             steps=config.steps, telemetry_interval_s=config.telemetry_interval_s)
    workflow = create_standalone_workflow(config)
    start = make_initial_params(config)
    if config.verbose:
        print_standalone_runtime_summary(workflow, start)
    result, summary = run_with_optional_telemetry(
        enabled=config.gpu_telemetry, interval_s=config.telemetry_interval_s,
        live=config.telemetry_live, label="Standalone backend telemetry",
        body=partial(_train, workflow, start))
    result.gpu_telemetry_summary = summary
    if config.verbose:
        print_result_summary(result, config)
    return result


__all__ = ["StandaloneResult", "StandaloneWorkflow", "create_standalone_workflow",
           "run_standalone"]
