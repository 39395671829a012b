"""torch.autograd binding of the MI355X RK4CN solver and its discrete adjoint.

``rk4cn(w0, solver, dt, nsteps, ckpt_every)`` maps a complex128 CUDA tensor
w0 ([Nx][Ny/2+1], FFTW units) to w_n after nsteps RK4CN steps on the solver's
GPU kernels; its backward is the hand-written discrete adjoint (ns2d_vjp_device:
checkpointed recompute of the RK stages, then the transposed RHS and stage
recursion), so torch losses on w_n (e.g. enstrophy) differentiate to w0.
torch is used for device memory and the autograd graph only.
"""
from __future__ import annotations

import torch


class _RK4CN(torch.autograd.Function):
    """CUDA tensors are handed to the solver by device pointer; CPU tensors
    are copied in and out (the computation runs on the GPU either way)."""

    @staticmethod
    def forward(ctx, w0, solver, dt, nsteps, ckpt_every):
        if w0.dtype != torch.complex128:
            raise TypeError("w0 must be complex128 ([Nx][Ny/2+1], FFTW units)")
        w0c = w0.detach().contiguous()
        if w0c.is_cuda:
            out = torch.empty_like(w0c)
            torch.cuda.current_stream().synchronize()
            solver.set_state_device(w0c.data_ptr())
            solver.spectral_step(dt, nsteps)
            solver.get_state_device(out.data_ptr())
        else:
            solver.set_state(w0c.numpy())
            solver.spectral_step(dt, nsteps)
            out = torch.from_numpy(solver.get_state())
        ctx.save_for_backward(w0c)
        ctx.solver, ctx.dt, ctx.nsteps, ctx.ckpt = solver, dt, nsteps, ckpt_every
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (w0,) = ctx.saved_tensors
        # torch's gradient of a real loss w.r.t. a complex tensor is dL/dRe + i dL/dIm,
        # the packing the adjoint uses
        g_T = grad_out.detach().contiguous()
        if w0.is_cuda:
            grad = torch.empty_like(w0)
            torch.cuda.current_stream().synchronize()
            ctx.solver.vjp_device(w0.data_ptr(), ctx.dt, ctx.nsteps, g_T.data_ptr(), grad.data_ptr(), ctx.ckpt)
        else:
            grad = torch.from_numpy(ctx.solver.vjp(w0.numpy(), ctx.dt, ctx.nsteps, g_T.numpy(), ctx.ckpt))
        return grad, None, None, None, None


def rk4cn(w0: torch.Tensor, solver, dt: float, nsteps: int, ckpt_every: int = 0) -> torch.Tensor:
    """w_n = Phi^nsteps(w0) with the GPU RK4CN step; differentiable w.r.t. w0."""
    return _RK4CN.apply(w0, solver, dt, nsteps, ckpt_every)


def enstrophy(w: torch.Tensor, Nx: int, Ny: int) -> torch.Tensor:
    """ComputeSystemMeasurables' total enstrophy (sys_msr.c:192,220,360) of w_hat."""
    import math
    c = torch.full((1, Ny // 2 + 1), 2.0, dtype=torch.float64, device=w.device)
    c = c.clone()
    c[0, 0] = 1.0
    c[0, -1] = 1.0
    a2 = w.real ** 2 + w.imag ** 2
    mask = torch.ones_like(a2)
    mask[0, 0] = 0.0
    return (c * a2 * mask).sum() * (4.0 * math.pi ** 2) * (0.5 / float(Nx * Ny) ** 2)
