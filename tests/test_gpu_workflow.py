"""End-to-end parity of the Python workflow on the GPU.

* The reference's stored 20-step gradient-descent energies (n = 8..24, L = 2,
  RunConfig defaults; tests/golden/gd20_final_energies.json) are reproduced by
  ``run(RunConfig(backend="standalone", ...))`` for its two strategies, and by
  dense_scan at n = 8.
* The one-step interface test of the reference
  (tests/test_training_workflow_interfaces.py:22-90) against the oracle.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from gpu_simulation_amd.ring_ising import (
    RingIsingAdjointBackend,
    RunConfig,
    StandaloneBackendConfig,
    run,
    run_standalone,
)

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "gd20_final_energies.json").read_text())
CFG = GOLDEN["run_config"]


def _cfg(n: int, strategy: str, **kw) -> RunConfig:
    base = dict(backend="standalone", num_qubits=n, layers=CFG["layers"], field=CFG["field"],
                steps=CFG["steps"], stepsize=CFG["stepsize"], seed=CFG["seed"],
                init_scale=CFG["init_scale"], gradient_strategy=strategy, verbose=False,
                show_progress=False)
    base.update(kw)
    return RunConfig(**base)


@pytest.mark.parametrize("row", GOLDEN["rows"], ids=lambda r: f"{r['method']}-{r['num_qubits']}q")
def test_engine_reproduces_reference_gd20_energy(row):
    n = row["num_qubits"]
    print(f"case {row['method']} n={n}", flush=True)
    result = run(_cfg(n, row["method"]))
    assert abs(result.final_energy - row["final_energy"]) <= 1e-10


def test_dense_scan_reproduces_reference_gd20_energy():
    row = next(r for r in GOLDEN["rows"] if r["num_qubits"] == 8)
    result = run_standalone(_cfg(8, "dense_scan"))
    assert abs(result.final_energy - row["final_energy"]) <= 1e-10


def test_one_step_interface_against_oracle(oracle):
    cfg = RunConfig(backend="standalone", num_qubits=4, layers=2, steps=1, report_steps=True,
                    show_progress=False, gradient_strategy="structured_adjoint", seed=123,
                    verbose=False)
    result = run_standalone(cfg)
    params = oracle.make_initial_params(4, 2, 123, 0.3)
    final_params, energies, final_energy = oracle.gd_loop(4, 2, 1.0, params, 1, 0.08)
    assert result.wall_s > 0.0
    assert len(result.step_metrics) == 1 and result.step_metrics[0].step == 1
    assert abs(result.step_metrics[0].energy - energies[0]) <= 1e-12
    assert abs(result.final_energy - final_energy) <= 1e-12
    assert np.max(np.abs(result.final_params - final_params)) <= 1e-12
    assert result.metadata["gradient_strategy"] == "structured_adjoint"


def test_backend_profile_and_alias(oracle):
    cfg = StandaloneBackendConfig(num_qubits=13, layers=2, field=0.9,
                                  gradient_strategy="inverse_walk_cuquantum")
    backend = RingIsingAdjointBackend(cfg)
    assert backend.gradient_strategy == "inverse_walk_cuQuantum"
    assert backend.uses_pennylane_gate_structure
    params = np.random.default_rng(3).uniform(0, 2 * np.pi, (2, 13, 2))
    out = backend.profile_energy_and_grad(params)
    e_ref, g_ref = oracle.energy_and_grad(13, 2, 0.9, params)
    assert abs(out["energy"] - e_ref) <= 1e-10
    assert np.max(np.abs(out["gradient"].reshape(-1) - g_ref)) <= 1e-10
    assert "backward_sweep_ms" in out["timings_ms"]
    with pytest.raises(ValueError):
        backend.energy(params[:, :12, :])


def test_cli_runs_the_standalone_workflow(oracle):
    from gpu_simulation_amd.ring_ising.cli import main

    result = main(["--qubits", "10", "--layers", "2", "--steps", "3", "--seed", "7"])
    assert result.final_params.shape == (2, 10, 2)
    e_ref, _ = oracle.energy_and_grad(10, 2, 1.0, result.final_params)
    assert abs(result.final_energy - e_ref) <= 1e-10


def test_stage_profiling_harness_writes_reference_columns(tmp_path):
    import csv

    from gpu_simulation_amd import profiling

    out = tmp_path / "stages.csv"
    rows = profiling.main(["--cases", "6x2", "14x2", "--repeats", "2", "--structured-widths", "2",
                           "--kernels", "--csv-out", str(out)])
    assert rows and out.exists()
    with out.open() as handle:
        read = list(csv.DictReader(handle))
    assert list(read[0].keys()) == profiling.CSV_FIELDS
    modes = {r["mode"] for r in read}
    assert {"inverse_walk_cuQuantum", "structured_w2", "dense_scan"} <= modes
    assert any(r["stage"].startswith("kernel:") for r in read)
