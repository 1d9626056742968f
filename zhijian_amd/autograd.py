"""torch.autograd bindings of the HIP residual and SSP-RK3 step.

The tensors are float64 on the handle's GPU, in the reference's AoS layout
[E][S][4] (contiguous).  The forward and the backward run on torch's current
stream through the library's device entry points; the backward is the
library's discrete adjoint (zj_rk3_step_vjp, zj_residual_vjp_device), so a
loss built from these functions differentiates through the reference
arithmetic of the path, with torch handling everything around it (initial
conditions, objectives, parameters of the initial state).
"""
from __future__ import annotations

import torch

from .solver import Solver


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _check(U: torch.Tensor, sol: Solver) -> torch.Tensor:
    if not U.is_cuda or U.dtype != torch.float64:
        raise TypeError("state tensors must be float64 on the GPU")
    if U.numel() != sol.n:
        raise ValueError(f"state has {U.numel()} values, the mesh needs {sol.n}")
    return U.contiguous()


class RK3Step(torch.autograd.Function):
    """U^{n+1} = SSP-RK3 step of U^n with step dt (FRSolver::advance,
    fr_solver.cpp:305-333, with the positivity limiter)."""

    @staticmethod
    def forward(ctx, U, dt, sol: Solver):
        U = _check(U, sol)
        dtv = float(dt)
        s = _stream()
        out = torch.empty_like(U)
        sol.set_state_device(U.data_ptr(), s)
        sol.rk3_step(dtv, s)
        sol.get_state_device(out.data_ptr(), s)
        ctx.save_for_backward(U)
        ctx.dt, ctx.sol, ctx.dt_is_tensor = dtv, sol, torch.is_tensor(dt)
        return out

    @staticmethod
    def backward(ctx, g):
        (U,) = ctx.saved_tensors
        sol = ctx.sol
        s = _stream()
        g = g.contiguous()
        out = torch.empty_like(U)
        sol.set_state_device(U.data_ptr(), s)
        sol.adjoint_set_device(g.data_ptr(), s)
        dtb = sol.rk3_step_vjp(ctx.dt, s)
        sol.adjoint_get_device(out.data_ptr(), s)
        gdt = torch.tensor(dtb, dtype=torch.float64, device=U.device) if ctx.dt_is_tensor else None
        return out, gdt, None


class Residual(torch.autograd.Function):
    """dU/dt = R(U) (FRSolver::computeResidual_GPU, fr_solver.cpp:335-608)."""

    @staticmethod
    def forward(ctx, U, sol: Solver):
        U = _check(U, sol)
        s = _stream()
        out = torch.empty_like(U)
        sol.set_state_device(U.data_ptr(), s)
        sol.residual_device(out.data_ptr(), s)
        ctx.save_for_backward(U)
        ctx.sol = sol
        return out

    @staticmethod
    def backward(ctx, g):
        (U,) = ctx.saved_tensors
        sol = ctx.sol
        s = _stream()
        g = g.contiguous()
        out = torch.empty_like(U)
        sol.set_state_device(U.data_ptr(), s)
        sol.residual_vjp_device(g.data_ptr(), out.data_ptr(), s)
        return out, None


def rk3_step(U: torch.Tensor, dt, sol: Solver) -> torch.Tensor:
    return RK3Step.apply(U, dt, sol)


def residual(U: torch.Tensor, sol: Solver) -> torch.Tensor:
    return Residual.apply(U, sol)
