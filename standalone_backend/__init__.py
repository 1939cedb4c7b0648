"""The reference's own import path for the engine: ``standalone_backend._cuda_backend``.

``RingIsingAdjointBackend._load_backend`` in the reference
(``ring_ising/backends/standalone/runtime.py:37-44``) imports
``standalone_backend._cuda_backend`` from the repository root.  This package
makes that name resolve to the pybind11 module built in
``gpu_simulation_amd/standalone_backend/`` (one module object under both
names), so reference code that imports it runs on this engine unchanged.
"""

from __future__ import annotations

import importlib
import sys

_cuda_backend = importlib.import_module("gpu_simulation_amd.standalone_backend._cuda_backend")
sys.modules[__name__ + "._cuda_backend"] = _cuda_backend

__all__ = ["_cuda_backend"]
