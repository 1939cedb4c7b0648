"""``torch.autograd.Function`` whose backward is the engine's adjoint sweep.

Forward runs the HIP forward sweep and ``lambda = H psi`` and returns the
energy; the final state and lambda stay resident in the engine's HBM buffers.
Backward runs the hand-written adjoint sweep on them: by default states are
recomputed by inverse gates rather than stored (optionally restored at
every-k-layer checkpoints, ``checkpoint_every``); ``schedule="stored"`` keeps
every forward pass output instead.  The incoming gradient scales the
parameter gradient on the device.  If another forward ran on the same
engine in between, backward first recomputes the forward for its own
parameters.  Parameters, energy and gradient never leave the GPU; torch only
provides the device memory and the current stream.
"""

from __future__ import annotations

import itertools

import torch

from .ring_ising.backends.standalone import RingIsingAdjointBackend, StandaloneBackendConfig

_generation = itertools.count(1)


class _EngineHandle:
    """A native engine plus the id of the forward whose state it holds."""

    def __init__(self, backend: RingIsingAdjointBackend) -> None:
        self.backend = backend
        self.native = backend.native
        self.resident_generation = 0
        self.param_shape = backend.config.param_shape


def _stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class RingIsingEnergyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params: torch.Tensor, handle: _EngineHandle) -> torch.Tensor:
        if not params.is_cuda or params.dtype != torch.float64:
            raise ValueError("params must be a float64 tensor on the GPU.")
        if tuple(params.shape) != handle.param_shape:
            raise ValueError(f"Expected params with shape {handle.param_shape}, "
                             f"got {tuple(params.shape)}.")
        flat = params.detach().contiguous().reshape(-1)
        energy = torch.empty(1, dtype=torch.float64, device=params.device)
        handle.native.forward_device(flat.data_ptr(), energy.data_ptr(),
                                     _stream_handle(params.device))
        gen = next(_generation)
        handle.resident_generation = gen
        ctx.handle = handle
        ctx.generation = gen
        ctx.save_for_backward(flat)
        return energy.reshape(())

    @staticmethod
    def backward(ctx, grad_energy: torch.Tensor):
        (flat,) = ctx.saved_tensors
        handle: _EngineHandle = ctx.handle
        stream = _stream_handle(flat.device)
        if handle.resident_generation != ctx.generation:
            scratch = torch.empty(1, dtype=torch.float64, device=flat.device)
            handle.native.forward_device(flat.data_ptr(), scratch.data_ptr(), stream)
        scale = grad_energy.detach().to(torch.float64).reshape(1).contiguous()
        grad = torch.empty_like(flat)
        handle.native.backward_device(flat.data_ptr(), scale.data_ptr(), grad.data_ptr(), stream)
        handle.resident_generation = 0
        return grad.reshape(handle.param_shape), None


class RingIsingEnergy(torch.nn.Module):
    """E(params) for params of shape (layers, n, 2) on the GPU, differentiable."""

    def __init__(self, num_qubits: int, layers: int, field: float = 1.0,
                 gradient_strategy: str = "structured_adjoint",
                 structured_rotation_chunk_width: int = 8,
                 checkpoint_every: int = 0, schedule: str = "inverse_walk") -> None:
        super().__init__()
        backend = RingIsingAdjointBackend(StandaloneBackendConfig(
            num_qubits=num_qubits, layers=layers, field=field,
            gradient_strategy=gradient_strategy,
            structured_rotation_chunk_width=structured_rotation_chunk_width))
        if gradient_strategy == "structured_adjoint":
            # "inverse_walk": states recomputed by inverse gates (optionally
            # restored at every-k-layer checkpoints); "stored": every forward
            # pass output is kept in HBM and read back.
            backend.native.set_schedule(schedule)
        if checkpoint_every:
            backend.native.set_checkpoint_interval(checkpoint_every)
        self.handle = _EngineHandle(backend)

    def forward(self, params: torch.Tensor) -> torch.Tensor:
        return RingIsingEnergyFunction.apply(params, self.handle)


def ring_ising_energy(params: torch.Tensor, module: RingIsingEnergy) -> torch.Tensor:
    return RingIsingEnergyFunction.apply(params, module.handle)
