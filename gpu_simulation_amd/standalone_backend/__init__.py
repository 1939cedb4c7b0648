"""Mount point of the pybind11 extension ``_cuda_backend`` (built in place by
``gpu_simulation_amd/build.py``), named as in the reference
(``standalone_backend/__init__.py``)."""

__all__ = ["_cuda_backend"]
