"""The allocation-free entry point rie_energy_and_grad_noalloc, through the C ABI.

It is checked against the allocating one-shot entry point rie_energy_and_grad
(structured_adjoint, same chunk width) at |dE|, |dgrad| <= 1e-10 on two calls
in a row, the second with a smaller circuit (the per-device scratch keeps the
first call's contents), plus an energy-only call and the size limits.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _lib():
    from gpu_simulation_amd import native

    lib = native.load_library()
    c = ctypes
    lib.rie_energy_and_grad_noalloc.argtypes = [
        c.c_size_t, c.c_size_t, c.c_double, c.c_void_p, c.c_size_t, c.c_size_t, c.c_void_p,
        c.c_void_p, c.c_void_p]
    lib.rie_energy_and_grad_noalloc.restype = c.c_int
    lib.rie_energy_and_grad.argtypes = [
        c.c_size_t, c.c_size_t, c.c_double, c.c_char_p, c.POINTER(c.c_double), c.c_size_t,
        c.c_int, c.c_int, c.c_size_t, c.POINTER(c.c_double), c.POINTER(c.c_double)]
    lib.rie_energy_and_grad.restype = c.c_int
    lib.rie_last_error.restype = c.c_char_p
    return lib


def _one_shot(lib, n, layers, field, params, width):
    flat = np.ascontiguousarray(params.reshape(-1), dtype=np.float64)
    energy = ctypes.c_double(0.0)
    grad = np.zeros(flat.size)
    dptr = ctypes.POINTER(ctypes.c_double)
    st = lib.rie_energy_and_grad(n, layers, field, b"structured_adjoint",
                                 flat.ctypes.data_as(dptr), flat.size, 1, 0, width,
                                 ctypes.byref(energy), grad.ctypes.data_as(dptr))
    assert st == 0, lib.rie_last_error()
    return energy.value, grad


def _no_alloc(lib, torch, n, layers, field, params, width, with_grad=True):
    p = torch.tensor(params.reshape(-1), dtype=torch.float64, device="cuda")
    e = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    g = torch.full_like(p, float("nan"))
    stream = torch.cuda.current_stream().cuda_stream
    st = lib.rie_energy_and_grad_noalloc(n, layers, field, p.data_ptr(), p.numel(), width,
                                         e.data_ptr(), g.data_ptr() if with_grad else None,
                                         stream)
    assert st == 0, lib.rie_last_error()
    torch.cuda.synchronize()
    return float(e.item()), g.cpu().numpy()


def test_no_alloc_matches_one_shot_on_two_calls():
    import torch

    lib = _lib()
    rng = np.random.default_rng(77)
    # a 14-qubit circuit (state scratch in the static array), then a smaller one
    for n, layers, field, width in [(14, 5, 0.9, 8), (9, 3, 1.3, 2)]:
        params = rng.uniform(0.0, 2.0 * np.pi, size=(layers, n, 2))
        print(f"case no-alloc n={n} L={layers} h={field} width={width}", flush=True)
        e_ref, g_ref = _one_shot(lib, n, layers, field, params, width)
        e, g = _no_alloc(lib, torch, n, layers, field, params, width)
        print(f"  |dE| = {abs(e - e_ref):.3e}  max|dgrad| = {np.max(np.abs(g - g_ref)):.3e}",
              flush=True)
        assert abs(e - e_ref) <= TOL
        assert np.max(np.abs(g - g_ref)) <= TOL
        e_only, g_untouched = _no_alloc(lib, torch, n, layers, field, params, width,
                                        with_grad=False)
        assert abs(e_only - e_ref) <= TOL
        assert np.isnan(g_untouched).all()


def test_no_alloc_rejects_sizes_beyond_its_scratch():
    import torch

    lib = _lib()
    p = torch.zeros(2 * 15 * 2, dtype=torch.float64, device="cuda")
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for n, layers, count in [(15, 2, 60), (4, 1025, 2 * 4 * 1025), (1, 2, 4), (4, 2, 15)]:
        st = lib.rie_energy_and_grad_noalloc(n, layers, 1.0, p.data_ptr(), count, 8,
                                             e.data_ptr(), None, stream)
        assert st == 1, (n, layers)  # RIE_INVALID_ARGUMENT, nothing launched
        assert lib.rie_last_error()
