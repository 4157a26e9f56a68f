"""Differentiable optical depth: torch.autograd.Function over the GPU forward
(fsamd_tau) and its hand-written adjoint (fsamd_tau_adjoint).

    tau = tau_autograd(plan, nbins, kernel, velfac, vel, temp, lines, dens)
    loss = ((torch.exp(-tau[0]) - f_target) ** 2).sum(); loss.backward()

gives d loss / d vel, temp and each line's density.  The discrete choices of
the forward (pixel windows set by the tail cut, the pixel a particle falls
in, the sub-sample count) are held fixed, so the gradient is that of the
piecewise-smooth function on the current piece.  The windows live in the
plan: another tau call on the same plan between forward and backward
overwrites them, and backward then raises instead of using them.  Gradients are accumulated in
FP64 and returned in the inputs' dtype (float32).  Neutral fraction or gas
density enter through dens = n * x_HI * const; build dens from them in torch
and autograd carries the chain rule.
"""
from __future__ import annotations

from typing import Sequence

import torch

from . import spectra_priv as SP


class _TauFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, nbins, kernel, velfac, lines, tautail, vel, temp, *dens):
        plan.keep_windows(True)
        outs = SP.tau(plan, nbins, kernel, velfac, vel.detach(), temp.detach(), lines,
                      [d.detach() for d in dens], tautail)
        ctx.plan = plan
        ctx.window_gen = plan.window_generation()
        ctx.args = (nbins, kernel, velfac, list(lines))
        ctx.save_for_backward(vel, temp, *dens)
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        vel, temp, *dens = ctx.saved_tensors
        nbins, kernel, velfac, lines = ctx.args
        plan = ctx.plan
        if plan.window_generation() != ctx.window_gen:
            raise RuntimeError("the plan's recorded pixel windows changed after this forward (another tau call on "
                               "the same plan); run backward before reusing the plan, or give each "
                               "differentiable call its own plan")
        gts = [(g if g is not None else torch.zeros((plan.nlos, nbins), device=vel.device)).to(torch.float64)
               .contiguous() for g in grads]
        gd, gT, gv = SP.tau_adjoint(plan, nbins, kernel, velfac, vel.detach(), temp.detach(), lines,
                                    [d.detach() for d in dens], gts)
        return (None, None, None, None, None, None, gv.to(vel.dtype), gT.to(temp.dtype),
                *[g.to(d.dtype) for g, d in zip(gd, dens)])


def tau_autograd(plan: SP.Plan, nbins: int, kernel: int, velfac: float, vel: torch.Tensor, temp: torch.Tensor,
                 lines: Sequence, dens: Sequence[torch.Tensor], tautail: float = 1e-7):
    """Optical depth of each line (float64 [nlos, nbins]) with gradients."""
    return list(_TauFunction.apply(plan, nbins, kernel, velfac, list(lines), tautail, vel, temp, *dens))
