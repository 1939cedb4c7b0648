"""Mount point of the pybind11 extension ``_cuda_backend`` (built in place by
``gpu_simulation_amd/build.py``), named as in the reference
(``standalone_backend/__init__.py``).

PyTorch, when installed, is imported before the extension: it ships its own
``libamdhip64`` under the soname the extension links, the process keeps the
first one loaded for both, and PyTorch's device only works with its own.
"""

try:
    import torch  # noqa: F401
except ImportError:
    pass

__all__ = ["_cuda_backend"]
