"""Write the oracle's energy and gradient for the 30-qubit x 2-layer workload.

The headline workload is ``structured_adjoint`` energy_and_grad at 30 qubits x
2 layers, J = h = 1, with parameters ``default_rng(5).uniform(0, 2*pi,
(2, 30, 2))`` (the same draw ``bench.py`` times).  This script evaluates it
with the CPU oracle's gate-level adjoint (``oracle/ring_ising_oracle.c``,
PennyLane gate order, two 16 GiB complex128 states) and writes
``tests/golden/c5_30q2l.json``: the parameters' generating rule, the energy,
all 120 gradient components and a digest of the parameters so a consumer can
check it regenerated the same draw.

The run needs about 33 GiB of host memory and several minutes on 8 cores; it
is made once in a container with that much memory and the JSON is committed.

Usage: python tests/golden/make_c5_fixture.py [--qubits 30] [--layers 2] [--seed 5]
"""

from __future__ import annotations

import argparse
import hashlib
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from oracle import oracle  # noqa: E402


def params_for(num_qubits: int, layers: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 2.0 * math.pi, size=(layers, num_qubits, 2)).reshape(-1)


def params_digest(params: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(params, dtype="<f8").tobytes()).hexdigest()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=2)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--field", type=float, default=1.0)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    n, layers = args.qubits, args.layers
    params = params_for(n, layers, args.seed)
    oracle.build()
    t0 = time.perf_counter()
    energy, grad = oracle.energy_and_grad(n, layers, args.field, params, form="gate")
    seconds = time.perf_counter() - t0
    out = Path(args.out) if args.out else Path(__file__).resolve().parent / f"c5_{n}q{layers}l.json"
    payload = {
        "description": ("CPU oracle gate-level adjoint (FP64, PennyLane gate order) of the "
                        "ring-Ising energy and gradient"),
        "num_qubits": n,
        "layers": layers,
        "field": args.field,
        "params_rule": f"numpy default_rng({args.seed}).uniform(0, 2*pi, ({layers}, {n}, 2)), flattened",
        "seed": args.seed,
        "params_sha256": params_digest(params),
        "energy": energy,
        "gradient": [float(g) for g in grad],
        "oracle_form": "gate",
        "oracle_seconds": round(seconds, 1),
        "oracle_threads": oracle.max_threads(),
    }
    out.write_text(json.dumps(payload, indent=1) + "\n")
    print(f"wrote {out}: E = {energy!r} in {seconds:.1f} s")


if __name__ == "__main__":
    main()
