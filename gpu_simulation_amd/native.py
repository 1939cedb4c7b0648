"""Locating and loading the built native code.

``librie.so`` (HIP kernels + C ABI) and the pybind11 module ``_cuda_backend``
are built in-tree by ``gpu_simulation_amd/build.py``.  There is no CPU
fallback: if they are missing or the GPU is absent, the calls fail loudly.
"""

from __future__ import annotations

import ctypes
import importlib

from . import build as _build

LIB_PATH = _build.LIB_PATH
EXT_PATH = _build.EXT_PATH

# Exported C ABI symbols (include/rie.h).
C_ABI_SYMBOLS = (
    "rie_last_error",
    "rie_abi_version",
    "rie_parse_strategy",
    "rie_device_count",
    "rie_engine_create",
    "rie_engine_destroy",
    "rie_engine_energy_and_grad",
    "rie_engine_timing_count",
    "rie_engine_timing",
    "rie_engine_benchmark_forward_graph",
    "rie_energy_and_grad",
    "rie_energy_and_grad_noalloc",
    "rie_engine_num_qubits",
    "rie_engine_num_layers",
    "rie_engine_strategy",
    "rie_engine_forward_device",
    "rie_engine_backward_device",
    "rie_engine_energy_and_grad_device",
    "rie_engine_set_checkpoint_interval",
    "rie_engine_inverse_walk_drift",
    "rie_engine_copy_state",
    "rie_engine_set_schedule",
    "rie_engine_schedule",
    "rie_engine_set_fused_hamiltonian",
    "rie_engine_set_pass_shapes",
    "rie_engine_pass_shapes",
    "rie_engine_fused_hamiltonian",
    "rie_engine_set_layer1_generation",
    "rie_engine_layer1_generation",
    "rie_engine_set_generated_start",
    "rie_engine_generated_start",
    "rie_engine_set_double_buffer",
    "rie_engine_double_buffer",
    "rie_engine_set_lambda_layouts",
    "rie_engine_lambda_layouts",
    "rie_engine_set_layer0_in_pass",
    "rie_engine_layer0_in_pass",
    "rie_engine_set_kernel_timing",
    "rie_engine_reset_kernel_stats",
    "rie_engine_kernel_stat_count",
    "rie_engine_kernel_stat",
    "rie_engine_set_tile_shapes",
    "rie_engine_tile_shapes",
    "rie_engine_autotune",
    "rie_batch_create",
    "rie_batch_destroy",
    "rie_batch_size",
    "rie_batch_num_params",
    "rie_batch_param_offsets",
    "rie_batch_strategies",
    "rie_batch_device_buffers",
    "rie_batch_set_params",
    "rie_batch_get_params",
    "rie_batch_get_results",
    "rie_batch_energy_and_grad",
    "rie_batch_run_device",
    "rie_batch_set_use_graph",
    "rie_batch_run_per_circuit",
    "rie_batch_optimize",
    "rie_batch_autotune",
    "rie_batch_set_tune_split",
    "rie_batch_set_config",
    "rie_batch_config",
    "rie_mlp_create",
    "rie_mlp_destroy",
    "rie_mlp_num_params",
    "rie_mlp_get_params",
    "rie_mlp_set_params",
    "rie_mlp_forward",
    "rie_mlp_loss_and_grad",
    "rie_mlp_train",
)


def ensure_built() -> None:
    """Build once if nothing is built yet (never silently substitutes code)."""
    if not _build.is_built():
        _build.build()


def _share_hip_runtime() -> None:
    """Load PyTorch's HIP runtime first when PyTorch is installed.

    PyTorch ships its own ``libamdhip64`` with the same soname as the ROCm one
    ``librie.so`` links; the process keeps whichever is loaded first for both.
    Loading PyTorch first keeps its runtime, so tensors it allocates and the
    engine's kernels share one HIP runtime and PyTorch's device stays usable.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library() -> ctypes.CDLL:
    ensure_built()
    _share_hip_runtime()
    return ctypes.CDLL(str(LIB_PATH))


def load_extension():
    """Import the pybind11 module ``_cuda_backend`` built from this tree (its
    package imports PyTorch first, as ``load_library`` does)."""
    ensure_built()
    return importlib.import_module("gpu_simulation_amd.standalone_backend._cuda_backend")
