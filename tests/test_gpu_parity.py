"""GPU parity: every strategy of the HIP engine against the FP64 CPU oracle.

The bar is |dE|, |dgrad| <= 1e-10 absolute in complex128 (the reference's own
tests use 1e-9 / 1e-8, tests/test_backends_parity.py:27-31).  Cases cover the
reference's test shapes (n = 2, 4, 6 at L = 2; n = 8 and 10 cross-strategy;
dense at 6q x 32L and 8q x 8L) plus every tile-pass geometry of the engine
(n = 13: one narrow high pass; 21: one full high pass; 22: two high passes).
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _params(seed: int, n: int, layers: int, scale: float | None = None) -> np.ndarray:
    rng = np.random.default_rng(seed)
    if scale is None:
        return rng.uniform(0.0, 2.0 * np.pi, size=(layers, n, 2))
    return scale * rng.standard_normal((layers, n, 2))


def _run(backend_module, n, layers, field, strategy, params, width=8, compute_gradient=True,
         schedule=None):
    eng = backend_module.RingIsingCudaBackend(n, layers, field, strategy, width, False)
    if schedule is not None:
        eng.set_schedule(schedule)
    out = eng.energy_and_grad(np.ascontiguousarray(params.reshape(-1)), compute_gradient)
    return out["energy"], np.asarray(out["gradient"])


def _check(oracle, backend_module, n, layers, field, strategy, seed, width=8, scale=None,
           schedule=None):
    params = _params(seed, n, layers, scale)
    print(f"case strategy={strategy} n={n} L={layers} h={field} seed={seed} schedule={schedule}",
          flush=True)
    e_ref, g_ref = oracle.energy_and_grad(n, layers, field, params)
    e, g = _run(backend_module, n, layers, field, strategy, params, width, schedule=schedule)
    assert abs(e - e_ref) <= TOL, (e, e_ref)
    assert g.shape == g_ref.shape
    assert np.max(np.abs(g - g_ref)) <= TOL, np.max(np.abs(g - g_ref))
    e_only, g_only = _run(backend_module, n, layers, field, strategy, params, width, False,
                          schedule=schedule)
    assert abs(e_only - e_ref) <= TOL
    assert g_only.size == 0


@pytest.mark.parametrize("strategy", ["structured_adjoint", "inverse_walk_cuQuantum", "dense_scan"])
@pytest.mark.parametrize("n,field", [(2, 0.7), (4, 1.0), (6, 1.1)])
@pytest.mark.parametrize("seed", [123, 321, 2026])
def test_reference_small_cases(oracle, backend_module, strategy, n, field, seed):
    # reference: tests/test_backends_parity.py:91-128 (params 0.2 N(0,1), L = 2)
    _check(oracle, backend_module, n, 2, field, strategy, seed, width=1, scale=0.2)


@pytest.mark.parametrize("n,layers", [(8, 1), (8, 3), (10, 2), (11, 4), (12, 3)])
def test_structured_whole_circuit_kernel(oracle, backend_module, n, layers):
    _check(oracle, backend_module, n, layers, 0.9, "structured_adjoint", 1000 + n)


@pytest.mark.parametrize("schedule", ["stored", "inverse_walk"])
@pytest.mark.parametrize("n,layers", [(13, 1), (13, 2), (14, 3), (16, 2), (21, 2), (22, 2),
                                      (22, 4), (25, 2)])
def test_structured_tile_passes(oracle, backend_module, n, layers, schedule):
    _check(oracle, backend_module, n, layers, 1.0, "structured_adjoint", 2000 + n,
           schedule=schedule)


@pytest.mark.parametrize("shapes", [(3, 3, 3), (4, 3, 3), (3, 4, 4)])
@pytest.mark.parametrize("schedule", ["stored", "inverse_walk"])
@pytest.mark.parametrize("n,layers", [(13, 2), (16, 3), (22, 2), (25, 2)])
def test_tile_register_shapes(oracle, backend_module, n, layers, schedule, shapes):
    params = _params(2100 + n, n, layers)
    print(f"case shapes={shapes} n={n} L={layers} schedule={schedule}", flush=True)
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 1.0, params)
    eng = backend_module.RingIsingCudaBackend(n, layers, 1.0, "structured_adjoint", 8, False)
    eng.set_schedule(schedule)
    eng.set_tile_shapes(*shapes)
    assert list(eng.tile_shapes) == list(shapes)
    out = eng.energy_and_grad(params.reshape(-1))
    assert abs(out["energy"] - e_ref) <= TOL
    assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL
    with pytest.raises(ValueError):
        eng.set_tile_shapes(5, 4, 4)
    # per pass kind: gen, fwd high, fwd low, H, bwd low, bwd high, layer-0 contraction
    mixed = [3, 4, 3, 4, 4, 3, 4]
    eng.set_pass_shapes(mixed)
    assert list(eng.pass_shapes) == mixed
    out = eng.energy_and_grad(params.reshape(-1))
    assert abs(out["energy"] - e_ref) <= TOL
    assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL
    with pytest.raises(ValueError):
        eng.set_pass_shapes([3, 4, 3])


@pytest.mark.parametrize("schedule", ["stored", "inverse_walk"])
@pytest.mark.parametrize("n,layers", [(13, 2), (14, 3), (16, 2), (21, 2), (22, 2), (22, 4),
                                      (25, 2)])
def test_fused_hamiltonian(oracle, backend_module, n, layers, schedule):
    # the fused evaluation (one scattered Hamiltonian pass, the other X terms
    # added by the last layer's adjoint passes) against the oracle and against
    # the per-group Hamiltonian passes; n = 13 / 21 / 22 / 25 give one narrow,
    # one full, two even and two uneven high groups
    params = _params(2300 + n, n, layers)
    print(f"case fused n={n} L={layers} schedule={schedule}", flush=True)
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 0.8, params)
    eng = backend_module.RingIsingCudaBackend(n, layers, 0.8, "structured_adjoint", 8, False)
    eng.set_schedule(schedule)
    assert eng.fused_hamiltonian is True
    fused = eng.energy_and_grad(params.reshape(-1))
    eng.set_fused_hamiltonian(False)
    assert eng.fused_hamiltonian is False
    plain = eng.energy_and_grad(params.reshape(-1))
    for out in (fused, plain):
        assert abs(out["energy"] - e_ref) <= TOL
        assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL
    assert abs(fused["energy"] - plain["energy"]) <= 1e-12
    assert np.max(np.abs(np.asarray(fused["gradient"]) - np.asarray(plain["gradient"]))) <= 1e-12
    eng.set_fused_hamiltonian(True)
    again = eng.energy_and_grad(params.reshape(-1))
    assert again["energy"] == fused["energy"]
    assert np.array_equal(np.asarray(again["gradient"]), np.asarray(fused["gradient"]))


def test_fused_hamiltonian_three_high_groups(backend_module):
    # 31 qubits: three high groups and four crossing wires (the scattered tile
    # holds index bits 0..7, 11, 18, 24, 30), beyond the oracle's reach; the
    # fused evaluation must equal the per-group Hamiltonian passes
    n, layers = 31, 2
    params = _params(2500, n, layers).reshape(-1)
    print(f"case fused n={n} L={layers}", flush=True)
    eng = backend_module.RingIsingCudaBackend(n, layers, 1.0, "structured_adjoint", 8, False)
    assert eng.fused_hamiltonian is True
    fused = eng.energy_and_grad(params)
    eng.set_fused_hamiltonian(0)
    plain = eng.energy_and_grad(params)
    assert abs(fused["energy"] - plain["energy"]) <= 1e-10
    assert np.max(np.abs(np.asarray(fused["gradient"]) - np.asarray(plain["gradient"]))) <= 1e-10
    assert abs(fused["energy"]) <= 2 * n


def test_fused_hamiltonian_device_path_and_scope(oracle, backend_module):
    import torch
    n, layers = 18, 3
    params = _params(2400, n, layers)
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 1.0, params)
    eng = backend_module.RingIsingCudaBackend(n, layers, 1.0, "structured_adjoint", 8, False)
    p = torch.tensor(params.reshape(-1), dtype=torch.float64, device="cuda")
    e = torch.empty(1, dtype=torch.float64, device="cuda")
    g = torch.empty_like(p)
    stream = torch.cuda.current_stream().cuda_stream
    eng.energy_and_grad_device(p.data_ptr(), e.data_ptr(), g.data_ptr(), stream)
    torch.cuda.synchronize()
    assert abs(float(e.item()) - e_ref) <= TOL
    assert np.max(np.abs(g.cpu().numpy() - g_ref)) <= TOL
    # one layer (product state + contraction) and the whole-circuit kernel keep
    # their own evaluation
    assert backend_module.RingIsingCudaBackend(16, 1, 1.0, "structured_adjoint", 8,
                                               False).fused_hamiltonian is False
    assert backend_module.RingIsingCudaBackend(10, 3, 1.0, "structured_adjoint", 8,
                                               False).fused_hamiltonian is False
    with pytest.raises(ValueError):
        eng.set_fused_hamiltonian(2)


@pytest.mark.parametrize("schedule", ["stored", "inverse_walk"])
@pytest.mark.parametrize("n,layers,shapes", [
    (13, 2, None), (14, 3, None), (16, 2, [4, 3, 3, 3, 3, 3, 3]), (21, 2, None),
    (22, 2, [3, 4, 4, 4, 3, 3, 4]), (22, 4, None), (25, 2, None)])
def test_layer1_generation(oracle, backend_module, n, layers, shapes, schedule):
    # layer 1 generated from the chain of 2x2 matrices (one forward write pass;
    # in the stored schedule the adjoint high passes of layer 1 generate the state)
    # against the oracle and against the pass-by-pass forward with stored
    # states; fused and unfused Hamiltonian; both register shapes of the
    # generator and of the generating adjoint passes
    params = _params(2700 + n, n, layers)
    print(f"case gen n={n} L={layers} shapes={shapes} schedule={schedule}", flush=True)
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 0.7, params)
    eng = backend_module.RingIsingCudaBackend(n, layers, 0.7, "structured_adjoint", 8, False)
    eng.set_schedule(schedule)
    if shapes is not None:
        eng.set_pass_shapes(shapes)
    flat = np.ascontiguousarray(params.reshape(-1))
    assert eng.layer1_generation is True
    outs = {}
    for gen in (1, 0):
        eng.set_layer1_generation(gen)
        assert eng.layer1_generation is bool(gen)
        for fused in (1, 0):
            eng.set_fused_hamiltonian(fused)
            outs[gen, fused] = eng.energy_and_grad(flat)
        outs[gen, "energy"] = eng.energy_and_grad(flat, False)["energy"]
    for key, out in outs.items():
        e = out if key[1] == "energy" else out["energy"]
        assert abs(e - e_ref) <= TOL, key
        if key[1] != "energy":
            assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL, key
    for fused in (1, 0):
        a, b = outs[1, fused], outs[0, fused]
        assert abs(a["energy"] - b["energy"]) <= 1e-12
        assert np.max(np.abs(np.asarray(a["gradient"]) - np.asarray(b["gradient"]))) <= 1e-12
    eng.set_layer1_generation(1)
    eng.set_fused_hamiltonian(1)
    again = eng.energy_and_grad(flat)
    assert again["energy"] == outs[1, 1]["energy"]
    assert np.array_equal(np.asarray(again["gradient"]), np.asarray(outs[1, 1]["gradient"]))


@pytest.mark.parametrize("kind", ["zero", "theta_pi", "quarter_turns"])
def test_layer1_generation_special_angles(oracle, backend_module, kind):
    # angles that put exact zeros and ones into the chain's site matrices
    n, layers = 16, 2
    params = np.zeros((layers, n, 2))
    if kind == "theta_pi":
        params[:, :, 0] = np.pi
    elif kind == "quarter_turns":
        params = (np.pi / 2) * np.random.default_rng(2900).integers(0, 4, size=(layers, n, 2))
    print(f"case gen special n={n} L={layers} kind={kind}", flush=True)
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 0.9, params)
    eng = backend_module.RingIsingCudaBackend(n, layers, 0.9, "structured_adjoint", 8, False)
    flat = np.ascontiguousarray(params.reshape(-1))
    for gen in (1, 0):
        eng.set_layer1_generation(gen)
        out = eng.energy_and_grad(flat)
        assert abs(out["energy"] - e_ref) <= TOL, gen
        assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL, gen


def test_layer1_generation_three_high_groups_and_scope(backend_module):
    # 31 qubits (three high groups, the largest size the chain tables cover):
    # generated against stored layer-1 states, beyond the oracle's reach
    n, layers = 31, 2
    params = _params(2800, n, layers).reshape(-1)
    print(f"case gen n={n} L={layers}", flush=True)
    eng = backend_module.RingIsingCudaBackend(n, layers, 1.0, "structured_adjoint", 8, False)
    assert eng.layer1_generation is True
    gen = eng.energy_and_grad(params)
    eng.set_layer1_generation(0)
    plain = eng.energy_and_grad(params)
    assert abs(gen["energy"] - plain["energy"]) <= 1e-10
    assert np.max(np.abs(np.asarray(gen["gradient"]) - np.asarray(plain["gradient"]))) <= 1e-10
    del eng
    # one layer and the whole-circuit kernel keep their own paths
    assert backend_module.RingIsingCudaBackend(16, 1, 1.0, "structured_adjoint", 8,
                                               False).layer1_generation is False
    assert backend_module.RingIsingCudaBackend(10, 3, 1.0, "structured_adjoint", 8,
                                               False).layer1_generation is False
    with pytest.raises(ValueError):
        backend_module.RingIsingCudaBackend(16, 2, 1.0, "structured_adjoint", 8,
                                            False).set_layer1_generation(2)


def test_engine_autotune_keeps_results(oracle, backend_module):
    n, layers = 18, 3
    params = _params(2200, n, layers)
    eng = backend_module.RingIsingCudaBackend(n, layers, 1.0, "structured_adjoint", 8, False)
    rep = eng.autotune(params.reshape(-1), 1)
    print("engine autotune", rep, flush=True)
    assert rep["tuned"] is True
    # per schedule: all kinds at shape 4, all at 3, and the per-kind choice
    # when it differs from both
    assert 4 <= len(rep["ms"]) <= 6
    assert list(eng.tile_shapes) == rep["tile_shapes"]
    assert list(eng.pass_shapes) == rep["pass_shapes"]
    assert eng.schedule == rep["schedule"]
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 1.0, params)
    out = eng.energy_and_grad(params.reshape(-1))
    assert abs(out["energy"] - e_ref) <= TOL
    assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL
    small = backend_module.RingIsingCudaBackend(8, 2, 1.0, "structured_adjoint", 8, False)
    assert small.autotune(_params(1, 8, 2).reshape(-1), 1) == {"tuned": False}


def test_schedule_selection(backend_module):
    small = backend_module.RingIsingCudaBackend(10, 3, 1.0, "structured_adjoint", 8, False)
    assert small.schedule == "inverse_walk"  # one fused kernel, nothing to store
    eng = backend_module.RingIsingCudaBackend(20, 3, 1.0, "structured_adjoint", 8, False)
    assert eng.schedule == "stored"  # fits easily
    eng.set_schedule("inverse_walk")
    assert eng.schedule == "inverse_walk"
    with pytest.raises(ValueError):
        eng.set_schedule("checkpoint")


@pytest.mark.parametrize("width", [1, 2, 3, 4, 8])
def test_structured_chunk_widths_agree(oracle, backend_module, width):
    # reference: tests/test_backends_parity.py:163-195
    _check(oracle, backend_module, 10, 2, 0.9, "structured_adjoint", 2030, width=width, scale=0.2)


@pytest.mark.parametrize("n,layers", [(8, 2), (13, 2)])
def test_gate_level_strategy(oracle, backend_module, n, layers):
    _check(oracle, backend_module, n, layers, 0.9, "inverse_walk_cuQuantum", 2028 + n)


@pytest.mark.parametrize("n,layers", [(3, 1), (5, 3), (6, 32), (7, 5), (8, 8), (8, 16),
                                      (4, 128), (7, 64), (8, 128)])
def test_dense_scan(oracle, backend_module, n, layers):
    # reference: tests/test_backends_parity.py:197-235 (6q x 32L, 8q x 8L);
    # 64- and 128-leaf product trees at the batch workload's depth
    _check(oracle, backend_module, n, layers, 1.0, "dense_scan", 2032 + n)


def test_profile_reports_reference_stage_names(backend_module):
    eng = backend_module.RingIsingCudaBackend(14, 2, 1.0, "structured_adjoint", 8, False)
    out = eng.energy_and_grad(_params(5, 14, 2).reshape(-1), True, True)
    names = set(out["timings_ms"])
    for key in ("build_ops_host_ms", "zero_state_ms", "forward_apply_ms", "hamiltonian_ms",
                "energy_dot_ms", "gradient_zero_ms", "backward_sweep_ms", "gradient_copy_ms"):
        assert key in names
    assert all(v >= 0.0 for v in out["timings_ms"].values())


def test_forward_graph_benchmark(backend_module):
    for n in (10, 14):
        eng = backend_module.RingIsingCudaBackend(n, 2, 1.0, "structured_adjoint", 8, False)
        res = eng.benchmark_structured_forward_graph(_params(6, n, 2).reshape(-1), 5)
        assert res["normal_forward_ms"] > 0.0 and res["graph_forward_ms"] > 0.0
        assert res["speedup"] > 0.0


def test_results_are_bitwise_reproducible(backend_module):
    params = _params(9, 16, 2).reshape(-1)
    eng = backend_module.RingIsingCudaBackend(16, 2, 1.0, "structured_adjoint", 8, False)
    a = eng.energy_and_grad(params)
    b = eng.energy_and_grad(params)
    assert a["energy"] == b["energy"]
    assert np.array_equal(a["gradient"], b["gradient"])


def test_final_state_matches_oracle(oracle, backend_module):
    n, layers = 14, 3
    params = _params(11, n, layers)
    eng = backend_module.RingIsingCudaBackend(n, layers, 1.0, "structured_adjoint", 8, False)
    eng.energy_and_grad(params.reshape(-1), False)
    psi = np.asarray(eng.final_state())
    ref = oracle.state(n, layers, params)
    assert np.max(np.abs(psi - ref)) <= 1e-12


@pytest.mark.parametrize("n,layers", [(10, 4), (20, 4)])
def test_inverse_walk_drift_and_checkpoints(oracle, backend_module, n, layers):
    params = _params(2024, n, layers)
    eng = backend_module.RingIsingCudaBackend(n, layers, 1.0, "structured_adjoint", 8, False)
    drift = eng.inverse_walk_drift(params.reshape(-1))
    assert 0.0 <= drift < 1e-11
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 1.0, params)
    eng.set_schedule("inverse_walk")  # checkpoints apply to the inverse walk
    eng.set_checkpoint_interval(2)
    out = eng.energy_and_grad(params.reshape(-1))
    assert abs(out["energy"] - e_ref) <= TOL
    assert np.max(np.abs(out["gradient"] - g_ref)) <= TOL


def test_module_level_energy_and_grad(oracle, backend_module):
    n, layers = 6, 2
    params = _params(7, n, layers)
    out = backend_module.energy_and_grad(n, layers, 1.0, "structured_adjoint", params.reshape(-1))
    e_ref, g_ref = oracle.energy_and_grad(n, layers, 1.0, params)
    assert abs(out["energy"] - e_ref) <= TOL
    assert np.max(np.abs(out["gradient"] - g_ref)) <= TOL
    with pytest.raises(ValueError):
        backend_module.energy_and_grad(n, layers, 1.0, "structured_adjoint", params)  # 3-D


@pytest.mark.parametrize("schedule", ["stored", "inverse_walk"])
@pytest.mark.parametrize("n,layers,shapes", [(13, 2, None), (16, 3, None), (22, 2, None),
                                              (22, 3, (3, 3, 3)), (25, 2, (4, 3, 4)),
                                              (31, 2, None)])
def test_double_buffer_is_result_identical(oracle, backend_module, n, layers, shapes, schedule):
    """double_buffer (next tile prefetched by the lambda-only adjoint passes)
    changes no arithmetic: bitwise-equal energy and gradient with it on and
    off, and the oracle's values where it fits."""
    params = np.ascontiguousarray(_params(7000 + n, n, layers).reshape(-1))
    outs = []
    for db in (False, True):
        print(f"case double_buffer={db} n={n} L={layers} schedule={schedule} shapes={shapes}",
              flush=True)
        eng = backend_module.RingIsingCudaBackend(n, layers, 1.0, "structured_adjoint", 8, db)
        eng.set_schedule(schedule)
        if shapes is not None:
            eng.set_tile_shapes(*shapes)
        assert eng.double_buffer is db
        out = eng.energy_and_grad(params)
        outs.append((out["energy"], np.asarray(out["gradient"])))
        eng.set_double_buffer(int(not db))
        assert eng.double_buffer is (not db)
        again = eng.energy_and_grad(params)
        assert again["energy"] == out["energy"]
        assert np.array_equal(np.asarray(again["gradient"]), outs[-1][1])
        del eng
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1])
    if n <= 25:
        e_ref, g_ref = oracle.energy_and_grad(n, layers, 1.0, params, form="structured")
        assert abs(outs[1][0] - e_ref) <= TOL
        assert np.max(np.abs(outs[1][1] - g_ref)) <= TOL


def test_double_buffer_rejects_bad_values(backend_module):
    eng = backend_module.RingIsingCudaBackend(13, 2, 1.0, "structured_adjoint", 8, False)
    with pytest.raises(ValueError):
        eng.set_double_buffer(2)


@pytest.mark.parametrize("n,shape", [(13, 4), (14, 3), (16, 4), (21, 4), (22, 3), (22, 4),
                                     (25, 4), (31, 4)])
def test_generated_start(oracle, backend_module, n, shape):
    """Two-layer gradient evaluations without forward or Hamiltonian passes:
    the low adjoint pass forms P^-1 H psi of its tile from generated
    amplitudes.  Against the oracle (n <= 25) and against the evaluation with
    the forward and the fused Hamiltonian pass (all n), at two fields."""
    for field, seed in [(1.0, 8100 + n), (0.37, 8200 + n)]:
        params = np.ascontiguousarray(_params(seed, n, 2).reshape(-1))
        eng = backend_module.RingIsingCudaBackend(n, 2, field, "structured_adjoint", 8, False)
        shapes = list(eng.pass_shapes)
        shapes[4] = shape  # the low adjoint pass's register shape
        eng.set_pass_shapes(shapes)
        print(f"case generated start n={n} shape={shape} h={field}", flush=True)
        assert eng.generated_start
        out = eng.energy_and_grad(params)
        with pytest.raises((RuntimeError, ValueError)):
            eng.final_state()  # no state is held after a generated-start evaluation
        eng.set_generated_start(0)
        assert not eng.generated_start
        ref = eng.energy_and_grad(params)
        assert abs(out["energy"] - ref["energy"]) <= TOL
        assert np.max(np.abs(np.asarray(out["gradient"]) - np.asarray(ref["gradient"]))) <= TOL
        if n <= 25:
            e_ref, g_ref = oracle.energy_and_grad(n, 2, field, params, form="structured")
            assert abs(out["energy"] - e_ref) <= TOL
            assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL
        e_only = eng.energy_and_grad(params, False)
        assert abs(e_only["energy"] - out["energy"]) <= TOL
        del eng


@pytest.mark.parametrize("n,r_low,r_high,r_l0", [(13, 4, 4, 4), (16, 3, 4, 3), (21, 4, 3, 4),
                                                  (21, 3, 4, 3), (22, 4, 4, 4), (25, 4, 4, 4),
                                                  (28, 3, 3, 3), (30, 4, 4, 4), (30, 3, 4, 3),
                                                  (31, 4, 4, 4)])
def test_lambda_layouts(oracle, backend_module, n, r_low, r_high, r_l0):
    """Generated-start evaluations with lambda stored between the passes in
    the layouts that make each high pass's tiles contiguous blocks: only the
    addresses change, so the results equal the natural order's bitwise; and
    against the oracle (n <= 25)."""
    params = np.ascontiguousarray(_params(8300 + n, n, 2).reshape(-1))
    eng = backend_module.RingIsingCudaBackend(n, 2, 0.8, "structured_adjoint", 8, False)
    shapes = list(eng.pass_shapes)
    shapes[4], shapes[5], shapes[6] = r_low, r_high, r_l0
    eng.set_pass_shapes(shapes)
    print(f"case lambda layouts n={n} r_low={r_low} r_high={r_high} r_l0={r_l0}", flush=True)
    assert eng.generated_start and eng.lambda_layouts
    out = eng.energy_and_grad(params)
    # the same pass order without layouts (where the order from the lowest
    # group up would lose the layouts, the engine runs from the top down; off,
    # it would switch)
    fused = eng.layer0_in_pass
    eng.set_lambda_layouts(0)
    eng.set_layer0_in_pass(1 if fused else 0)
    assert not eng.lambda_layouts and eng.layer0_in_pass == fused
    ref = eng.energy_and_grad(params)
    assert out["energy"] == ref["energy"]
    assert np.array_equal(np.asarray(out["gradient"]), np.asarray(ref["gradient"]))
    if n <= 25:
        e_ref, g_ref = oracle.energy_and_grad(n, 2, 0.8, params, form="structured")
        assert abs(out["energy"] - e_ref) <= TOL
        assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL
    with pytest.raises(ValueError):
        eng.set_lambda_layouts(2)
    del eng


def test_lambda_layouts_scope(backend_module):
    # 19 and 20 qubits: the single high group [n - 9, n) holds some of the low
    # tile's register bits 8..11, which the low pass's layout then splits, so
    # lambda stays in the natural order; without the generated start the
    # layouts are not used
    for n in (19, 20):
        eng = backend_module.RingIsingCudaBackend(n, 2, 1.0, "structured_adjoint", 8, False)
        assert eng.generated_start and not eng.lambda_layouts
        del eng
    eng = backend_module.RingIsingCudaBackend(16, 2, 1.0, "structured_adjoint", 8, False)
    assert eng.lambda_layouts
    eng.set_generated_start(0)
    assert not eng.lambda_layouts


@pytest.mark.parametrize("n,layouts", [(13, 1), (16, 1), (19, 1), (21, 0), (22, 1), (25, 1),
                                       (26, 0), (30, 1)])
def test_layer0_in_pass(oracle, backend_module, n, layouts):
    """Two-layer evaluations whose last high pass (the top group, high passes
    from the lowest group up) contracts layer 0 instead of storing lambda:
    against the separate contraction (the high passes from the top group
    down) at 1e-12, and the oracle (n <= 25), at two fields, one, two and
    three high groups, with and without lambda layouts, special angles
    included."""
    for field, seed in [(1.0, 8400 + n), (0.61, 8500 + n)]:
        params = _params(seed, n, 2)
        if field != 1.0:
            # zeros in the layer-0 factors of wires 3, n - 1 and a fixed wire
            params[0, 3, 0] = np.pi
            params[0, n - 1, 0] = 0.0
            params[0, min(5, n - 10), 0] = np.pi
        params = np.ascontiguousarray(params.reshape(-1))
        eng = backend_module.RingIsingCudaBackend(n, 2, field, "structured_adjoint", 8, False)
        eng.set_lambda_layouts(layouts)
        print(f"case layer0 in pass n={n} layouts={layouts} h={field}", flush=True)
        assert eng.generated_start and eng.layer0_in_pass
        out = eng.energy_and_grad(params)
        eng.set_layer0_in_pass(0)
        assert not eng.layer0_in_pass
        ref = eng.energy_and_grad(params)
        assert abs(out["energy"] - ref["energy"]) <= 1e-12
        assert np.max(np.abs(np.asarray(out["gradient"]) - np.asarray(ref["gradient"]))) <= 1e-12
        if n <= 25:
            e_ref, g_ref = oracle.energy_and_grad(n, 2, field, params, form="structured")
            assert abs(out["energy"] - e_ref) <= TOL
            assert np.max(np.abs(np.asarray(out["gradient"]) - g_ref)) <= TOL
        del eng


def test_layer0_in_pass_scope(backend_module):
    # on by default where the top group's pass can take it; off with the
    # 3-register shape for the high passes, and at 28 qubits, where the
    # passes from the lowest group up would lose the lambda layouts
    eng = backend_module.RingIsingCudaBackend(16, 2, 1.0, "structured_adjoint", 8, False)
    assert eng.layer0_in_pass
    shapes = list(eng.pass_shapes)
    shapes[5] = 3
    eng.set_pass_shapes(shapes)
    assert not eng.layer0_in_pass
    eng = backend_module.RingIsingCudaBackend(28, 2, 1.0, "structured_adjoint", 8, False)
    assert eng.lambda_layouts and not eng.layer0_in_pass
    eng.set_lambda_layouts(0)
    assert eng.layer0_in_pass
    del eng
    eng = backend_module.RingIsingCudaBackend(16, 3, 1.0, "structured_adjoint", 8, False)
    assert not eng.layer0_in_pass
    with pytest.raises(ValueError):
        eng.set_layer0_in_pass(2)


def test_generated_start_scope(backend_module):
    # three layers, unfused H or no layer-1 generation: the forward runs as before
    eng = backend_module.RingIsingCudaBackend(16, 3, 1.0, "structured_adjoint", 8, False)
    assert not eng.generated_start
    eng = backend_module.RingIsingCudaBackend(16, 2, 1.0, "structured_adjoint", 8, False)
    assert eng.generated_start
    eng.set_fused_hamiltonian(0)
    assert not eng.generated_start
    eng.set_fused_hamiltonian(1)
    eng.set_layer1_generation(0)
    assert not eng.generated_start
    with pytest.raises(ValueError):
        eng.set_generated_start(3)
