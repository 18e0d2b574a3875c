"""torch.autograd.Function over the hand-written adjoint (gradient/inverse harness).

    J = SHLoss.apply(psi0, eps, q0, cfg, nsteps, ckpt_every, target_amp)

psi0 is a float64 tensor of shape (nx, ny, nz); eps and q0 are 0-d float64
tensors.  Forward runs nsteps periodic CNAB2 steps on the GPU (libsmectic) and
returns J = 1/2 sum (psi_T - target_amp cos(q0 z))^2; backward returns dJ/dpsi0,
dJ/deps and dJ/dq0 from smx_adjoint_grad (checkpointed recomputation).  The
grid (lx, ly, lz) is held fixed while q0 varies.
"""
import copy

import numpy as np
import torch

from .solver import Solver


class SHLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi0, eps, q0, cfg, nsteps, ckpt_every=50, target_amp=0.5):
        c = copy.copy(cfg)
        c.eps = float(eps)
        c.q0 = float(q0)
        with Solver(c) as s:
            J, d, de, dq = s.adjoint_grad(psi0.detach().cpu().numpy(), nsteps, ckpt_every, target_amp)
        ctx.save_for_backward(torch.from_numpy(d), torch.tensor(de, dtype=torch.float64),
                              torch.tensor(dq, dtype=torch.float64))
        ctx.device = psi0.device
        return torch.tensor(J, dtype=torch.float64, device=psi0.device)

    @staticmethod
    def backward(ctx, gJ):
        d, de, dq = ctx.saved_tensors
        g = gJ.detach().cpu()
        return ((g * d).to(ctx.device), (g * de).to(ctx.device), (g * dq).to(ctx.device), None, None, None, None)
