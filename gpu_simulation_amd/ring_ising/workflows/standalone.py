"""Standalone (HIP) training workflow (``ring_ising/workflows/standalone.py:90-167``).

Creates the backend, draws the initial parameters, runs the plain
gradient-descent loop with one ``energy_and_grad`` per step and evaluates the
energy of the final parameters.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from ..backends.standalone import RingIsingAdjointBackend, StandaloneBackendConfig
from ..config import RunConfig
from ..training import (
    StepEvaluation,
    TrainingRunResult,
    run_gradient_descent_loop,
    run_with_optional_telemetry,
    validate_common_run_args,
)
from .common import apply_gradient_step, make_initial_params, print_problem_summary, \
    print_result_summary


@dataclass(frozen=True)
class StandaloneWorkflow:
    config: RunConfig
    backend_config: StandaloneBackendConfig
    backend: RingIsingAdjointBackend


StandaloneResult = TrainingRunResult


def _backend_config(config: RunConfig) -> StandaloneBackendConfig:
    return StandaloneBackendConfig(
        num_qubits=config.num_qubits,
        layers=config.layers,
        field=config.field,
        gradient_strategy=config.gradient_strategy,
        structured_rotation_chunk_width=config.effective_structured_rotation_chunk_width,
        double_buffer=config.double_buffer,
    )


def create_standalone_workflow(config: RunConfig) -> StandaloneWorkflow:
    backend_config = _backend_config(config)
    return StandaloneWorkflow(config=config, backend_config=backend_config,
                              backend=RingIsingAdjointBackend(backend_config))


def print_standalone_runtime_summary(workflow: StandaloneWorkflow, params: np.ndarray) -> None:
    config = workflow.config
    print("Runtime:")
    print("  Backend: Standalone HIP (MI355X)")
    print(f"  Gradient strategy: {workflow.backend.gradient_strategy}")
    if workflow.backend.gradient_strategy == "structured_adjoint":
        print("  Structured rotation chunk width: "
              f"{workflow.backend_config.effective_structured_rotation_chunk_width}")
    if workflow.backend.uses_pennylane_gate_structure:
        print("  PennyLane-style gate structure: True")
    else:
        print("  Dense-scan fused gate structure: True")
    print(f"  Estimated gradient workspace: {workflow.backend.estimated_workspace_gib:.2f} GiB")
    print()
    print_problem_summary(config, params)
    print("Backend details:")
    print(f"  Initial energy: {workflow.backend.energy(params):.10f}")
    print()


def run_standalone(config: RunConfig) -> TrainingRunResult:
    if config.backend != "standalone":
        raise ValueError(f"run_standalone expected backend='standalone', got {config.backend!r}.")
    validate_common_run_args(num_qubits=config.num_qubits, layers=config.layers,
                             steps=config.steps, telemetry_interval_s=config.telemetry_interval_s)
    workflow = create_standalone_workflow(config)
    params = make_initial_params(config)
    if config.verbose:
        print_standalone_runtime_summary(workflow, params)

    def step_fn(current: np.ndarray) -> StepEvaluation:
        energy, grad = workflow.backend.energy_and_grad(current)
        return StepEvaluation(energy=float(energy), grad=np.asarray(grad, dtype=np.float64))

    def body() -> TrainingRunResult:
        start = time.perf_counter()
        loop = run_gradient_descent_loop(
            params,
            steps=config.steps,
            stepsize=config.stepsize,
            step_fn=step_fn,
            apply_gradient_step=apply_gradient_step,
            verbose=config.verbose,
            show_progress=config.show_progress and not config.report_steps,
            report_steps=config.report_steps,
            format_step=lambda m: (f"step={m.step:03d} energy={m.energy:.10f} "
                                   f"grad_norm={m.grad_norm:.10f} "
                                   f"step_ms={1000.0 * m.step_wall_s:.3f}"),
        )
        final_energy = workflow.backend.energy(loop.final_params)
        return TrainingRunResult(
            backend_label="standalone_hip",
            final_params=np.asarray(loop.final_params, dtype=np.float64),
            final_energy=float(final_energy),
            step_metrics=loop.step_metrics,
            wall_s=time.perf_counter() - start,
            metadata={
                "gradient_strategy": workflow.backend.gradient_strategy,
                "structured_rotation_chunk_width":
                    workflow.backend_config.effective_structured_rotation_chunk_width,
                "estimated_workspace_gib": workflow.backend.estimated_workspace_gib,
                "pennylane_gate_structure": workflow.backend.uses_pennylane_gate_structure,
            },
        )

    result, summary = run_with_optional_telemetry(
        enabled=config.gpu_telemetry, interval_s=config.telemetry_interval_s,
        live=config.telemetry_live, label="Standalone backend telemetry", body=body)
    result.gpu_telemetry_summary = summary
    if config.verbose:
        print_result_summary(result, config)
    return result


__all__ = ["StandaloneResult", "StandaloneWorkflow", "create_standalone_workflow",
           "run_standalone"]
