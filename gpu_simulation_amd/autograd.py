"""``torch.autograd.Function`` whose backward is the engine's adjoint sweep.

Forward runs the HIP forward sweep and ``lambda = H psi`` and returns the
energy; the final state and lambda stay resident in the engine's HBM buffers.
Backward runs the hand-written adjoint sweep on them: by default states are
recomputed by inverse gates rather than stored (optionally restored at
every-k-layer checkpoints, ``checkpoint_every``, either always or -- with
``drift_tol`` -- only once the measured inverse-walk drift of the module's
parameters exceeds that tolerance); ``schedule="stored"`` keeps every forward
pass output instead.  The incoming gradient scales the
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
    """E(params) for params of shape (layers, n, 2) on the GPU, differentiable.

    Checkpoints (inverse-walk schedule): ``checkpoint_every=k`` stores the state
    at every k-th layer boundary in the forward sweep and restores it in the
    backward sweep instead of walking through it.  With ``drift_tol`` set they
    start switched off; the inverse-walk drift ||walked state - |0...0>||_2 of
    the parameters is measured on the first call and then every
    ``drift_check_every`` calls (0: first call only), and once it exceeds
    ``drift_tol`` checkpoints every ``checkpoint_every`` layers (2 when not
    given) are switched on for the rest of the module's life.  ``last_drift``
    and ``checkpoints_active`` report the state of that mode.
    """

    def __init__(self, num_qubits: int, layers: int, field: float = 1.0,
                 gradient_strategy: str = "structured_adjoint",
                 structured_rotation_chunk_width: int = 8,
                 checkpoint_every: int = 0, schedule: str = "inverse_walk",
                 drift_tol: float | None = None, drift_check_every: int = 0) -> None:
        super().__init__()
        if drift_tol is not None:
            if gradient_strategy != "structured_adjoint" or schedule != "inverse_walk":
                raise ValueError("drift-triggered checkpoints need structured_adjoint with the "
                                 "inverse_walk schedule.")
            if not drift_tol >= 0.0 or drift_check_every < 0:
                raise ValueError("drift_tol must be >= 0 and drift_check_every >= 0.")
        backend = RingIsingAdjointBackend(StandaloneBackendConfig(
            num_qubits=num_qubits, layers=layers, field=field,
            gradient_strategy=gradient_strategy,
            structured_rotation_chunk_width=structured_rotation_chunk_width))
        if gradient_strategy == "structured_adjoint":
            # "inverse_walk": states recomputed by inverse gates (optionally
            # restored at every-k-layer checkpoints); "stored": every forward
            # pass output is kept in HBM and read back.
            backend.native.set_schedule(schedule)
        self.drift_tol = drift_tol
        self.drift_check_every = drift_check_every
        self.checkpoint_every = checkpoint_every or (2 if drift_tol is not None else 0)
        self.checkpoints_active = False
        self.last_drift: float | None = None
        self._calls = 0
        if checkpoint_every and drift_tol is None:
            backend.native.set_checkpoint_interval(checkpoint_every)
            self.checkpoints_active = True
        self.handle = _EngineHandle(backend)

    def _watch_drift(self, params: torch.Tensor) -> None:
        due = self._calls == 0 or (self.drift_check_every > 0 and
                                   self._calls % self.drift_check_every == 0)
        self._calls += 1
        if self.checkpoints_active or not due:
            return
        host = params.detach().to("cpu", torch.float64).contiguous().numpy().reshape(-1)
        self.last_drift = float(self.handle.native.inverse_walk_drift(host))
        # the measurement overwrote the engine's buffers: a pending backward
        # re-runs its own forward first
        self.handle.resident_generation = 0
        if self.last_drift > self.drift_tol:
            self.handle.native.set_checkpoint_interval(self.checkpoint_every)
            self.checkpoints_active = True

    def forward(self, params: torch.Tensor) -> torch.Tensor:
        if self.drift_tol is not None:
            self._watch_drift(params)
        return RingIsingEnergyFunction.apply(params, self.handle)


def ring_ising_energy(params: torch.Tensor, module: RingIsingEnergy) -> torch.Tensor:
    return RingIsingEnergyFunction.apply(params, module.handle)
