"""torch.autograd binding of the VSVB energy: E(coeff) with dE/dcoeff from the
adjoint pass of the HIP library (one call gives both).

This is the in-process form of what vtools does by running the ``valence``
binary per energy (vtools/iotools.py:487-495, vtools/vtools.py:1558-1566): the
LCAO weights become a differentiable tensor, so any torch optimizer can drive
them, replacing the per-weight energy differences of the reference optimizer
(demgs_opt, src/valence.F90:351-521).

CUDA tensors go through the device-pointer entry on torch's current stream (no
host staging); CPU tensors go through the host entry.  There is no CPU fallback:
without the HIP library every call raises.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import lib
from .system import VsvbSystem


class VsvbEnergy(torch.autograd.Function):
    """E = energy(coeff) for the wave function layout of ``context``; backward
    returns grad_out * dE/dcoeff (the adjoint gradient saved by forward)."""

    @staticmethod
    def forward(ctx, coeff: torch.Tensor, context: lib.Context):
        if coeff.dtype != torch.float64 or coeff.dim() != 1 or coeff.numel() != context.sys.totlen:
            raise ValueError(f"coeff must be a float64 vector of {context.sys.totlen} weights")
        c = coeff.detach().contiguous()
        if c.is_cuda:
            e = torch.empty((), dtype=torch.float64, device=c.device)
            g = torch.empty_like(c)
            stream = torch.cuda.current_stream(c.device).cuda_stream
            context.energy_grad_device(c.data_ptr(), e.data_ptr(), g.data_ptr(), stream)
        else:
            ev, gv = context.energy_grad(c.numpy())
            e = torch.tensor(ev, dtype=torch.float64)
            g = torch.from_numpy(gv)
        ctx.save_for_backward(g)
        return e

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None


def energy(context: lib.Context, coeff: torch.Tensor) -> torch.Tensor:
    """Differentiable VSVB energy (Eh) of ``coeff`` (float64, totlen weights)."""
    return VsvbEnergy.apply(coeff, context)


def optimize_weights(s: VsvbSystem, steps: int = 20, lr: float = 1.0, device: Optional[str] = None,
                     history: Optional[list] = None) -> np.ndarray:
    """Minimize the energy over the LCAO weights with L-BFGS driven by the adjoint
    gradient (one library call per energy + gradient).  Returns the optimized
    weights; ``history`` receives the energy after each step."""
    dev = torch.device(device) if device else torch.device("cuda" if torch.cuda.is_available() else "cpu")
    with lib.Context(s) as ctx:
        c = torch.tensor(s.coeff, dtype=torch.float64, device=dev, requires_grad=True)
        opt = torch.optim.LBFGS([c], lr=lr, max_iter=1, history_size=20, line_search_fn="strong_wolfe")

        def closure():
            opt.zero_grad()
            e = energy(ctx, c)
            e.backward()
            return e

        for _ in range(steps):
            e = opt.step(closure)
            if history is not None:
                history.append(float(e))
        return c.detach().cpu().numpy()
