"""Many independent ring-Ising circuits evaluated together on one GPU.

The reference evaluates one circuit per ``RingIsingCudaBackend.energy_and_grad``
call (``cpp/ising_cuda_backend.cu:429-442``).  :class:`RingIsingBatch` wraps the
extension's batch object (C ABI ``rie_batch_*``): circuits of different qubit
counts (2..14), layer counts and fields share one persistent launch of the
whole-circuit kernel (structured_adjoint) and grouped FP64-MFMA product-tree
launches (dense_scan, ``num_qubits <= dense_max_qubits``), optionally replayed
from a HIP graph, with gradient-descent or Adam steps on the device.

Parameters of circuit ``i`` have the reference's per-circuit layout, shape
``(L_i, n_i, 2)`` with ``[..., 0] = theta_RY`` and ``[..., 1] = phi_RZ``; the
batch packs them one after the other.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import native


@dataclass(frozen=True)
class CircuitSpec:
    num_qubits: int
    layers: int
    field: float


def uneven_mix(count: int, seed: int = 3, qubits: tuple[int, int] = (4, 14),
               layer_choices: Sequence[int] = (8, 32, 128),
               field_range: tuple[float, float] = (0.2, 2.0)):
    """A reproducible mix of circuits: circuit i draws from ``default_rng([seed, i])``
    n uniform in ``qubits`` (inclusive), L from ``layer_choices``, h uniform in
    ``field_range``, then its parameters uniform in [0, 2 pi)."""
    specs, params = [], []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        n = int(rng.integers(qubits[0], qubits[1] + 1))
        layers = int(rng.choice(np.asarray(layer_choices)))
        h = float(rng.uniform(*field_range))
        specs.append(CircuitSpec(n, layers, h))
        params.append(rng.uniform(0.0, 2.0 * np.pi, size=(layers, n, 2)))
    return specs, params


class RingIsingBatch:
    """A fixed set of circuits with device-resident packed parameters."""

    def __init__(self, specs: Sequence[CircuitSpec], dense_max_qubits: int = 8):
        ext = native.load_extension()
        self.specs = list(specs)
        self._b = ext.RingIsingBatch([s.num_qubits for s in self.specs],
                                     [s.layers for s in self.specs],
                                     [float(s.field) for s in self.specs], int(dense_max_qubits))
        self.offsets = np.asarray(self._b.param_offsets(), dtype=np.int64)

    @property
    def strategies(self) -> list[str]:
        """Each circuit's strategy; autotune() may move circuits between them."""
        return list(self._b.strategies())

    @property
    def native(self):
        return self._b

    def __len__(self) -> int:
        return len(self.specs)

    @property
    def num_params(self) -> int:
        return int(self._b.num_params)

    def pack(self, params: Sequence[np.ndarray]) -> np.ndarray:
        if len(params) != len(self.specs):
            raise ValueError("one parameter array per circuit is required.")
        flat = np.empty(self.num_params, dtype=np.float64)
        for i, (s, p) in enumerate(zip(self.specs, params)):
            arr = np.asarray(p, dtype=np.float64)
            if arr.size != 2 * s.num_qubits * s.layers:
                raise ValueError(f"circuit {i}: expected {2 * s.num_qubits * s.layers} parameters.")
            flat[self.offsets[i]:self.offsets[i + 1]] = arr.reshape(-1)
        return flat

    def unpack(self, flat: np.ndarray) -> list[np.ndarray]:
        flat = np.asarray(flat)
        return [flat[self.offsets[i]:self.offsets[i + 1]].reshape(s.layers, s.num_qubits, 2)
                for i, s in enumerate(self.specs)]

    def energy_and_grad(self, params: Sequence[np.ndarray], compute_gradient: bool = True):
        energies, grad = self._b.energy_and_grad(self.pack(params), compute_gradient)
        return np.asarray(energies), (self.unpack(grad) if compute_gradient else None)

    def results(self):
        """Energies and per-circuit gradients of the last evaluation on the device."""
        energies, grad = self._b.results()
        return np.asarray(energies), self.unpack(grad)

    def set_params(self, params: Sequence[np.ndarray]) -> None:
        self._b.set_params(self.pack(params))

    def get_params(self) -> list[np.ndarray]:
        return self.unpack(self._b.get_params())

    def optimize(self, optimizer: str, lr: float, steps: int, use_graph: bool = True,
                 stream: int = 0) -> np.ndarray:
        """Run ``steps`` optimizer steps from the current parameters; returns the
        (steps, count) energies before each step."""
        return np.asarray(self._b.optimize(optimizer, float(lr), int(steps), bool(use_graph),
                                           int(stream)))

    def autotune(self, repeats: int = 3, choose_split: bool = True) -> dict:
        """Tune kernel shapes, and with choose_split which circuits with
        num_qubits <= dense_max_qubits take dense_scan, on this batch."""
        self._b.set_tune_split(int(bool(choose_split)))
        return self._b.autotune(int(repeats))

    def config(self) -> dict:
        return self._b.config()


def main(argv=None) -> None:
    """Evaluate an uneven mix a few times (a fixed workload for profilers):

        python -m gpu_simulation_amd.batch --count 2048 --evals 3 [--dense-only] [--dense-max N]
    """
    import argparse
    import time

    p = argparse.ArgumentParser(description="evaluate an uneven batch of circuits")
    p.add_argument("--count", type=int, default=2048)
    p.add_argument("--seed", type=int, default=3)
    p.add_argument("--evals", type=int, default=3)
    p.add_argument("--dense-only", action="store_true", help="keep only the n <= 8 circuits")
    p.add_argument("--dense-max", type=int, default=8,
                   help="circuits with n <= this take dense_scan (the autotuner's choice on the "
                        "2048-circuit mix is 4)")
    p.add_argument("--threads", type=int, default=256)
    p.add_argument("--chunks", default="3,3,3,3,3,3,3,3,3,2,2,3,4,4,4",
                   help="rotation chunk width per qubit count 0..14")
    p.add_argument("--dense-tile", type=int, default=64)
    p.add_argument("--grid", type=int, default=128, help="structured persistent workgroups")
    p.add_argument("--autotune", action="store_true", help="autotune instead of the fixed config")
    args = p.parse_args(argv)
    specs, params = uneven_mix(args.count, seed=args.seed)
    if args.dense_only:
        keep = [i for i, s in enumerate(specs) if s.num_qubits <= 8]
        specs, params = [specs[i] for i in keep], [params[i] for i in keep]
    batch = RingIsingBatch(specs, dense_max_qubits=args.dense_max)
    batch.set_params(params)
    if args.autotune:
        batch.autotune(repeats=2)
    else:
        batch.native.set_config(args.threads, [int(c) for c in args.chunks.split(",")],
                                args.dense_tile, 0 if args.dense_only else args.grid)
    batch.results()  # synchronise
    t0 = time.perf_counter()
    for _ in range(args.evals):
        batch.native.run_device(True, 0)
    energies, _ = batch.results()
    dt = (time.perf_counter() - t0) / max(args.evals, 1)
    print(f"{len(specs)} circuits, {dt * 1e3:.3f} ms per evaluation of all, "
          f"mean energy {float(np.mean(energies)):.6f}", flush=True)


if __name__ == "__main__":
    main()
