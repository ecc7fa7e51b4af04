"""torch.autograd.Function over the checkpointed adjoint of the plain step.

    psi_K = rollout(psi0, theta)   theta = (eps, alpha, beta, gamma)

Forward runs `nsteps` plain steps on the GPU stepper with checkpoints every
`ckpt_every` steps; backward runs the discrete adjoint for the incoming
cotangent dL/dpsi_K and returns dL/dpsi0 and dL/dtheta.  The reference has
no gradients; this is new functionality on top of the plain-model step
(no-adv/cosNoAdvTrackMult.cpp:729-805).

GPU tensors stay on the device: psi0, psi_K, the cotangent and dL/dpsi0
cross the C ABI as device FP64 buffers (smad_forward_tape_dev,
smad_adjoint_dev), and contexts are reused from a pool keyed by model, grid
and precision (smad_set_params installs theta), so a training loop neither
copies fields through the host nor rebuilds a context per call.  CPU tensors
go through the host-buffer entry points (the work still runs on the GPU).
"""
from __future__ import annotations

import dataclasses
import threading

import numpy as np
import torch

from . import Params, Stepper

_POOL: dict = {}
_POOL_LOCK = threading.Lock()


def _key(p: Params):
    return (p.model, p.nx, p.ny, p.nz, p.precision)


def _acquire(p: Params, psi0_host=None) -> Stepper:
    """A free context for p's model, grid and precision with p's
    coefficients, from the pool or new."""
    with _POOL_LOCK:
        free = _POOL.get(_key(p), [])
        s = free.pop() if free else None
    if s is None:
        return Stepper(p, psi0_host if psi0_host is not None else np.zeros(p.shape))
    s.set_params(p)
    return s


def _release(s: Stepper):
    with _POOL_LOCK:
        _POOL.setdefault(_key(s.params), []).append(s)


class PlainRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi0, theta, params: Params, nsteps: int, ckpt_every: int):
        th = theta.detach().cpu().double().numpy()
        p = dataclasses.replace(params, eps=float(th[0]), alpha=float(th[1]),
                                beta=float(th[2]), gamma=float(th[3]))
        ctx.meta = (psi0.device, psi0.dtype, theta.device, theta.dtype)
        if psi0.is_cuda:
            x = psi0.detach().to(torch.float64).contiguous()
            out = torch.empty_like(x)
            s = _acquire(p)
            torch.cuda.current_stream(x.device).synchronize()
            s.forward_tape_dev(x.data_ptr(), int(nsteps), int(ckpt_every), out.data_ptr())
            ctx.stepper = s
            return out.to(psi0.dtype)
        x = psi0.detach().cpu().double().numpy()
        s = _acquire(p, x)
        s.set_field(0, x)  # the pre-loop on psi0
        s.forward_tape(int(nsteps), int(ckpt_every))
        ctx.stepper = s
        return torch.from_numpy(s.fetch()).to(dtype=psi0.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        s = ctx.stepper
        dev0, dt0, devt, dtt = ctx.meta
        try:
            if dev0.type == "cuda":
                g = grad_out.detach().to(device=dev0, dtype=torch.float64).contiguous()
                d0 = torch.empty_like(g)
                torch.cuda.current_stream(dev0).synchronize()
                dp = s.adjoint_dev(g.data_ptr(), d0.data_ptr())
                d0 = d0.to(dt0)
            else:
                d0n, dp = s.adjoint(grad_out.detach().cpu().double().numpy())
                d0 = torch.from_numpy(d0n).to(dtype=dt0)
        finally:
            ctx.stepper = None
            _release(s)
        return d0, torch.from_numpy(np.asarray(dp)).to(device=devt, dtype=dtt), None, None, None


def rollout(psi0: torch.Tensor, theta: torch.Tensor, params: Params, nsteps: int,
            ckpt_every: int = 20) -> torch.Tensor:
    """Differentiable K-step plain rollout; theta = (eps, alpha, beta, gamma)."""
    return PlainRollout.apply(psi0, theta, params, nsteps, ckpt_every)
