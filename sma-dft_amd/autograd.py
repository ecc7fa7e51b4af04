"""torch.autograd.Function over the checkpointed adjoint of the plain step.

    psi_K = rollout(psi0, theta)   theta = (eps, alpha, beta, gamma)

Forward runs `nsteps` plain steps on the GPU stepper with checkpoints every
`ckpt_every` steps (smad_forward_tape); backward runs the discrete adjoint
(smad_adjoint) for the incoming cotangent dL/dpsi_K and returns dL/dpsi0 and
dL/dtheta.  Tensors cross the C ABI as host FP64 arrays; results come back
on the input's device and dtype.  The reference has no gradients; this is
new functionality on top of the plain-model step
(no-adv/cosNoAdvTrackMult.cpp:729-805).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import Params, Stepper


class PlainRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi0, theta, params: Params, nsteps: int, ckpt_every: int):
        th = theta.detach().cpu().double().numpy()
        p = dataclasses.replace(params, eps=float(th[0]), alpha=float(th[1]),
                                beta=float(th[2]), gamma=float(th[3]))
        s = Stepper(p, psi0.detach().cpu().double().numpy())
        s.forward_tape(int(nsteps), int(ckpt_every))
        psik = s.fetch()
        ctx.stepper = s
        ctx.meta = (psi0.device, psi0.dtype, theta.device, theta.dtype)
        return torch.from_numpy(psik).to(device=psi0.device, dtype=psi0.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        s = ctx.stepper
        dev0, dt0, devt, dtt = ctx.meta
        d0, dp = s.adjoint(grad_out.detach().cpu().double().numpy())
        s.close()
        ctx.stepper = None
        return (torch.from_numpy(d0).to(device=dev0, dtype=dt0),
                torch.from_numpy(np.asarray(dp)).to(device=devt, dtype=dtt),
                None, None, None)


def rollout(psi0: torch.Tensor, theta: torch.Tensor, params: Params, nsteps: int,
            ckpt_every: int = 20) -> torch.Tensor:
    """Differentiable K-step plain rollout; theta = (eps, alpha, beta, gamma)."""
    return PlainRollout.apply(psi0, theta, params, nsteps, ckpt_every)
