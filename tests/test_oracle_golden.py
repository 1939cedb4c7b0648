"""The oracle reproduces the reference's stored end-to-end results.

Fixture: ``tests/golden/gd20_final_energies.json`` (final energies of the
reference's 20-step gradient-descent runs at RunConfig defaults, L = 2; both
its structured and its cuStateVec gate-level strategy).  Matching them pins
parameter initialisation, ansatz, Hamiltonian, gradient and update together.
"""

from __future__ import annotations

import json
from pathlib import Path

import pytest

GOLDEN = json.loads((Path(__file__).parent / "golden" / "gd20_final_energies.json").read_text())
CFG = GOLDEN["run_config"]
TOL = 1e-11


def _rows(max_n: int):
    return [r for r in GOLDEN["rows"] if r["num_qubits"] <= max_n]


@pytest.mark.parametrize("row", _rows(18), ids=lambda r: f"{r['method']}-{r['num_qubits']}q")
@pytest.mark.parametrize("form", ["gate", "structured"])
def test_oracle_matches_reference_gd20_energy(oracle, row, form):
    n = row["num_qubits"]
    params = oracle.make_initial_params(n, CFG["layers"], CFG["seed"], CFG["init_scale"])
    _, _, final = oracle.gd_loop(n, CFG["layers"], CFG["field"], params, CFG["steps"],
                                 CFG["stepsize"], form=form)
    assert abs(final - row["final_energy"]) <= TOL


def test_golden_fixture_is_self_consistent():
    # the reference's two strategies and the sibling CSV's gate-level rows agree
    by_n = {}
    for r in GOLDEN["rows"]:
        by_n.setdefault(r["num_qubits"], []).append(r["final_energy"])
    for r in GOLDEN["sibling_gate_level_rows"]:
        by_n[r["num_qubits"]].append(r["final_energy"])
    assert sorted(by_n) == list(range(8, 25))
    for n, values in by_n.items():
        assert max(values) - min(values) < 1e-12 * max(1.0, n), (n, values)
