"""The C ABI library loads and exports every declared symbol (CPU only).

No compute runs here: the checks cover symbol export, the strategy parser
and the error path of engine creation when no GPU is visible.
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from gpu_simulation_amd import native

HEADER = Path(__file__).resolve().parents[1] / "include" / "rie.h"


def _declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rie_[a-z0-9_]+)\s*\(", text)))


def test_header_and_loader_list_agree():
    assert _declared_functions() == sorted(native.C_ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = native.load_library()
    for name in _declared_functions():
        assert hasattr(lib, name), name


def test_strategy_parser_and_error_reporting():
    lib = native.load_library()
    lib.rie_parse_strategy.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
    lib.rie_parse_strategy.restype = ctypes.c_int
    lib.rie_last_error.restype = ctypes.c_char_p
    out = ctypes.c_int(-1)
    for name, value in [(b"inverse_walk_cuQuantum", 0), (b"inverse_walk_cuquantum", 0),
                        (b"structured_adjoint", 1), (b"dense_scan", 2)]:
        assert lib.rie_parse_strategy(name, ctypes.byref(out)) == 0
        assert out.value == value
    assert lib.rie_parse_strategy(b"checkpoint", ctypes.byref(out)) == 1  # RIE_INVALID_ARGUMENT
    assert b"Unknown gradient strategy" in lib.rie_last_error()
    assert lib.rie_abi_version() == 1


def test_extension_maps_errors_like_the_reference():
    ext = native.load_extension()
    with pytest.raises(ValueError):
        ext.RingIsingCudaBackend(4, 2, 1.0, "ryrz_fused", 8, False)
    if ext.device_count() == 0:
        with pytest.raises(RuntimeError):
            ext.RingIsingCudaBackend(4, 2, 1.0, "structured_adjoint", 8, False)


def test_batch_validates_arguments_before_touching_the_device():
    ext = native.load_extension()
    with pytest.raises(ValueError):
        ext.RingIsingBatch([4, 15], [2, 2], [1.0, 1.0])  # 15 qubits: beyond the batch kernel
    with pytest.raises(ValueError):
        ext.RingIsingBatch([4, 5], [2, 0], [1.0, 1.0])  # zero layers
    with pytest.raises(ValueError):
        ext.RingIsingBatch([4, 5], [2], [1.0, 1.0])  # ragged descriptions
    with pytest.raises(ValueError):
        ext.RingIsingBatch([], [], [])  # empty batch
    with pytest.raises(ValueError):
        ext.RingIsingBatch([4], [2], [1.0], 9)  # dense_scan stops at 8 qubits
    if ext.device_count() == 0:
        with pytest.raises(RuntimeError):
            ext.RingIsingBatch([4, 5], [2, 2], [1.0, 1.0])


def test_uneven_mix_is_reproducible():
    from gpu_simulation_amd.batch import uneven_mix

    a_specs, a_params = uneven_mix(16, seed=3)
    b_specs, b_params = uneven_mix(16, seed=3)
    assert a_specs == b_specs
    assert all(np.array_equal(x, y) for x, y in zip(a_params, b_params))
    assert all(4 <= s.num_qubits <= 14 and s.layers in (8, 32, 128) and 0.2 <= s.field <= 2.0
               for s in a_specs)


def test_reference_import_path_is_the_built_module():
    """``standalone_backend._cuda_backend`` -- the name the reference's
    ``RingIsingAdjointBackend._load_backend`` imports -- is this build's module."""
    import importlib

    ext = native.load_extension()
    alias = importlib.import_module("standalone_backend._cuda_backend")
    assert alias is ext
    assert hasattr(alias, "RingIsingCudaBackend") and hasattr(alias, "energy_and_grad")
