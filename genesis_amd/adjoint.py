"""Reverse-mode gradients of trajectory observables as a torch op.

``TrajectoryLoss.apply(theta, universe, keep, n_steps, ckpt_every, win_from,
gr, g0)`` runs ``gs_adjoint_forward`` (n_steps velocity-Verlet ticks with
checkpoints) and returns the loss as a 0-d float64 tensor; ``backward``
calls ``gs_adjoint_backward`` (hand-written adjoint kernels: Hessian-vector
products and parameter VJPs of the switched LJ + Coulomb law, recomputation
from the checkpoints).  ``theta`` is the 30-vector of per-species sigma,
epsilon (kelvin) and charge.  The state starts from what is resident in the
universe; call ``universe.upload(state)`` to reset it between evaluations.
"""
from __future__ import annotations

import numpy as np
import torch

from .params import Physics


def theta_of(physics: Physics) -> torch.Tensor:
    return torch.tensor(list(physics.sigma) + list(physics.epsilon_k) + list(physics.charge), dtype=torch.float64)


class TrajectoryLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, universe, keep, n_steps, ckpt_every, win_from, gr=None, g0=None):
        th = theta.detach().to("cpu", torch.float64).numpy()
        loss, _ = universe.adjoint_forward(th, keep, int(n_steps), int(ckpt_every), int(win_from), gr, g0)
        ctx.universe = universe
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g = torch.from_numpy(ctx.universe.adjoint_backward())
        return grad_out * g, None, None, None, None, None, None, None


def trajectory_loss(theta, universe, keep, n_steps, ckpt_every, win_from, gr=None, g0=None):
    return TrajectoryLoss.apply(theta, universe, np.asarray(keep, dtype=np.uint8), n_steps, ckpt_every, win_from,
                                gr, g0)
