"""Stage and kernel profiling of energy_and_grad (the reference's
``benchmarks/profile_baseline_stages.py:67-186`` harness, same cases, flags and CSV
columns), plus per-kernel HIP-event timings of the tile passes.

    python -m gpu_simulation_amd.profiling --cases 12x8 20x4 --csv-out stages.csv

Stage names are the reference's ``timings_ms`` keys
(``cpp/ising_cuda_structured_adjoint_modes.inc:317-370``).  With ``--kernels`` the
structured rows are followed by one row per kernel (``stage`` = kernel name, median
milliseconds per evaluation) from the engine's own launch timing.
"""

from __future__ import annotations

import argparse
import csv
import gc
import re
import statistics as st
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .ring_ising.backends.standalone import RingIsingAdjointBackend, StandaloneBackendConfig
from .ring_ising.params import make_initial_params_array

CSV_FIELDS = ["num_qubits", "layers", "mode", "gradient_strategy",
              "structured_rotation_chunk_width", "stage", "median_ms", "total_stage_ms"]


@dataclass(frozen=True)
class BenchmarkCase:
    num_qubits: int
    layers: int


_CASE = re.compile(r"(\d+)[xX](\d+)")


def parse_case(text: str) -> BenchmarkCase:
    """``"QxL"`` (e.g. ``12x16``) -> ``BenchmarkCase(Q, L)``; argparse error otherwise."""
    match = _CASE.fullmatch(text.strip())
    if match is None:
        raise argparse.ArgumentTypeError(f"Invalid case {text!r}. Expected format like 12x16.")
    return BenchmarkCase(*map(int, match.groups()))


def parse_args(argv=None) -> argparse.Namespace:
    p = argparse.ArgumentParser(description=__doc__,
                                formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--cases", nargs="+", type=parse_case,
                   default=[parse_case(c) for c in ("12x8", "12x32", "12x128", "16x4", "16x8",
                                                    "16x16", "20x2", "20x4", "20x8")])
    p.add_argument("--field", type=float, default=1.0)
    p.add_argument("--seed", type=int, default=7)
    p.add_argument("--init-scale", type=float, default=0.3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--repeats", type=int, default=20)
    p.add_argument("--structured-widths", "--mode2-widths", dest="structured_widths", nargs="+",
                   type=int, default=[1, 2, 3, 4, 8])
    p.add_argument("--kernels", action="store_true",
                   help="also record per-kernel launch timings of structured_adjoint")
    p.add_argument("--csv-out", type=Path, default=None)
    return p.parse_args(argv)


def stage_profile_summary(backend: RingIsingAdjointBackend, params: np.ndarray, *, warmup: int,
                          repeats: int) -> tuple[float, dict[str, float]]:
    """Median total and per-stage medians of ``timings_ms`` over ``repeats`` profiled calls.

    The stage timings come from HIP events inside the engine; Python's cyclic
    collector is paused over the timed calls so that no collection lands
    between an evaluation's stages.
    """
    for _ in range(warmup):
        backend.profile_energy_and_grad(params)
    runs: list[dict[str, float]] = []
    collecting = gc.isenabled()
    gc.disable()
    try:
        for _ in range(repeats):
            timings = backend.profile_energy_and_grad(params)["timings_ms"]
            runs.append({name: float(ms) for name, ms in timings.items()})
    finally:
        if collecting:
            gc.enable()
    stages = sorted({name for run in runs for name in run})
    per_stage = {name: st.median([run[name] for run in runs if name in run]) for name in stages}
    return st.median([sum(run.values()) for run in runs]), per_stage


def kernel_profile(backend: RingIsingAdjointBackend, params: np.ndarray, *,
                   repeats: int) -> dict[str, float]:
    """Milliseconds per evaluation of every kernel, from HIP events on the launch stream."""
    eng = backend.native
    flat = np.ascontiguousarray(params.reshape(-1))
    eng.energy_and_grad(flat)
    eng.reset_kernel_stats()
    eng.set_kernel_timing(True)
    for _ in range(repeats):
        eng.energy_and_grad(flat)
    eng.set_kernel_timing(False)
    return {k["name"]: k["total_ms"] / repeats for k in eng.kernel_stats() if k["launches"]}


def profile(args: argparse.Namespace) -> list[dict[str, object]]:
    rows: list[dict[str, object]] = []
    for case in args.cases:
        params = make_initial_params_array(num_qubits=case.num_qubits, layers=case.layers,
                                           seed=args.seed, init_scale=args.init_scale)
        print(f"{case.num_qubits} qubits x {case.layers} layers", flush=True)
        specs = [("inverse_walk_cuQuantum", "inverse_walk_cuQuantum", 1)]
        specs += [(f"structured_w{w}", "structured_adjoint", w) for w in args.structured_widths]
        if case.num_qubits <= 8:
            specs.append(("dense_scan", "dense_scan", 1))
        for label, mode, width in specs:
            backend = RingIsingAdjointBackend(StandaloneBackendConfig(
                num_qubits=case.num_qubits, layers=case.layers, field=args.field,
                gradient_strategy=mode, structured_rotation_chunk_width=width))
            total, stages = stage_profile_summary(backend, params, warmup=args.warmup,
                                                  repeats=args.repeats)
            print(f"  {label:<22} total_stage_ms={total:.3f}", flush=True)
            for name, value in stages.items():
                print(f"    {name:<24} {value:9.3f}")
                rows.append({"num_qubits": case.num_qubits, "layers": case.layers, "mode": label,
                             "gradient_strategy": mode, "structured_rotation_chunk_width": width,
                             "stage": name, "median_ms": value, "total_stage_ms": total})
            if args.kernels and mode == "structured_adjoint":
                for name, value in sorted(kernel_profile(backend, params,
                                                         repeats=args.repeats).items()):
                    print(f"    kernel {name:<24} {value:9.3f}")
                    rows.append({"num_qubits": case.num_qubits, "layers": case.layers,
                                 "mode": label + "_kernels", "gradient_strategy": mode,
                                 "structured_rotation_chunk_width": width,
                                 "stage": "kernel:" + name, "median_ms": value,
                                 "total_stage_ms": total})
    return rows


def main(argv=None) -> list[dict[str, object]]:
    args = parse_args(argv)
    print("Standalone stage profiling (MI355X)")
    rows = profile(args)
    if args.csv_out is not None:
        args.csv_out.parent.mkdir(parents=True, exist_ok=True)
        with args.csv_out.open("w", encoding="utf-8", newline="") as handle:
            writer = csv.DictWriter(handle, fieldnames=CSV_FIELDS)
            writer.writeheader()
            writer.writerows(rows)
        print(f"CSV written to {args.csv_out}")
    return rows


if __name__ == "__main__":
    main()
