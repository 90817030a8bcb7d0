"""torch.autograd bridge for the discrete adjoint of the GPU time stepper.

`rollout(solver, u0, dt, nsteps, snapshots)` returns the state after `nsteps` explicit RK steps of
the solver's scheme from the modal coefficients `u0` (reference layout E x N_b x NV, float64);
its backward pass runs the checkpointed discrete adjoint of the library (sdg_adjoint_forward /
sdg_adjoint_backward), so any differentiable torch objective of the final state can be
back-propagated to the initial state.

The state crosses to NumPy (CPU) for set_state / get_state; a CUDA gradient of the final state is
handed to the reverse sweep in device memory (sdg_adjoint_backward_device), a CPU one through the
host entry point. The stepping and the adjoint run on the GPU either way.
"""
from __future__ import annotations

import numpy as np
import torch


class DGRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0: torch.Tensor, solver, dt: float, nsteps: int, snapshots: int):
        solver.set_state(u0.detach().cpu().numpy().astype(np.float64))
        solver.adjoint_forward(dt, nsteps, snapshots)
        un = solver.get_state()
        ctx.solver = solver
        return torch.from_numpy(un).to(dtype=u0.dtype, device=u0.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        if grad_out.is_cuda:  # device-resident reverse sweep: no host round trip of the adjoints
            lam = grad_out.detach().to(torch.float64).contiguous()
            return ctx.solver.adjoint_backward_device(lam).to(grad_out.dtype), None, None, None, None
        lam0 = ctx.solver.adjoint_backward(grad_out.detach().cpu().numpy().astype(np.float64))
        return torch.from_numpy(lam0).to(dtype=grad_out.dtype, device=grad_out.device), None, None, None, None


def rollout(solver, u0: torch.Tensor, dt: float, nsteps: int, snapshots: int | None = None) -> torch.Tensor:
    """Differentiable N-step rollout; `snapshots` defaults to one checkpoint per step."""
    return DGRollout.apply(u0, solver, dt, nsteps, nsteps if snapshots is None else snapshots)
