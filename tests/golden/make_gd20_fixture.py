"""Extract the reference's stored 20-step gradient-descent energies into a fixture.

Reads (as data) the reference CSV
``ring_cx_pipeline_vs_original_structed_vs_inversecuQuantum/
originstructed_vs_cuquantum_8to24q_2layers_summary_merged.csv`` and writes
``tests/golden/gd20_final_energies.json``.  The run settings are the
reference's ``RunConfig`` defaults (``ring_ising/config.py:26-53``): field 1.0,
seed 7, init_scale 0.3, stepsize 0.08, 20 steps; the CSV fixes layers = 2.

The sibling CSV ``ring_cnot_pipeline_vs_cuquantum_8to24q_2layers_merged.csv``
is not used for structured rows: its structured energies disagree with both
gate-level columns (e.g. n = 8: -8.790185 vs -8.829768).  Its gate-level rows
are recorded as a second, independent copy.
"""

from __future__ import annotations

import csv
import json
from pathlib import Path

REF = Path("/root/reference/ring_cx_pipeline_vs_original_structed_vs_inversecuQuantum")
GOOD = REF / "originstructed_vs_cuquantum_8to24q_2layers_summary_merged.csv"
SIBLING = REF / "ring_cnot_pipeline_vs_cuquantum_8to24q_2layers_merged.csv"
OUT = Path(__file__).resolve().parent / "gd20_final_energies.json"


def main() -> None:
    rows = []
    with GOOD.open() as handle:
        for line_no, row in enumerate(csv.DictReader(handle), start=2):
            rows.append({
                "csv_line": line_no,
                "method": row["method"],
                "num_qubits": int(row["num_qubits"]),
                "layers": int(row["layers"]),
                "steps": int(row["steps"]),
                "final_energy": float(row["final_energy"]),
                "avg_step_ms_rtx6000ada": float(row["avg_step_ms"]),
            })
    sibling_gate = []
    with SIBLING.open() as handle:
        for line_no, row in enumerate(csv.DictReader(handle), start=2):
            if row["method"] == "inverse_walk_cuQuantum":
                sibling_gate.append({
                    "csv_line": line_no,
                    "num_qubits": int(row["num_qubits"]),
                    "final_energy": float(row["final_energy"]),
                })
    payload = {
        "source": str(GOOD.relative_to(REF.parent)),
        "run_config": {"field": 1.0, "seed": 7, "init_scale": 0.3, "stepsize": 0.08,
                       "steps": 20, "layers": 2},
        "rows": rows,
        "sibling_gate_level_rows": sibling_gate,
    }
    OUT.write_text(json.dumps(payload, indent=1) + "\n")
    print(f"wrote {OUT} ({len(rows)} rows)")


if __name__ == "__main__":
    main()
