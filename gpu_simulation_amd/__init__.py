"""MI355X-native ring-Ising statevector engine.

* ``include/rie.h`` / ``lib/librie.so``: the C ABI and the gfx950 HIP kernels;
* ``standalone_backend._cuda_backend``: the pybind11 module with the
  reference extension's Python surface;
* ``ring_ising``: the reference's Python workflow surface (``RunConfig``,
  ``run``, ``RingIsingAdjointBackend`` ...);
* ``autograd``: ``torch.autograd.Function`` wrapper whose backward is the
  hand-written adjoint sweep.
"""

__version__ = "0.1.0"
