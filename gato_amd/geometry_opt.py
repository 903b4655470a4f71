"""Alternating SCF + Adam geometry optimisation (reference: geometry_opt.py:28-78).

Adam follows Kingma & Ba (2015) with the defaults optax.adam documents
(b1 = 0.9, b2 = 0.999, eps = 1e-8, eps_root = 0), which is what the
reference calls at geometry_opt.py:62.  The update runs on the (K, 3) host
position tensor; the electronic solves and forces run on the GPU.
"""
from __future__ import annotations

from typing import Callable, TypeVar

import torch

from .geometry import Nuclei, recenter

Point = TypeVar("Point")


class Adam:
    def __init__(self, learning_rate: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.lr, self.b1, self.b2, self.eps = learning_rate, b1, b2, eps
        self.m = None
        self.v = None
        self.t = 0

    def update(self, grad: torch.Tensor) -> torch.Tensor:
        g = grad.detach().to(torch.float64).cpu()
        if self.m is None:
            self.m = torch.zeros_like(g)
            self.v = torch.zeros_like(g)
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return -self.lr * mh / (torch.sqrt(vh) + self.eps)


def optimize_geometry_alternating(initial_nuclei: Nuclei, *,
                                  solve_at_geometry: Callable[[Nuclei, "Point | None"], Point],
                                  bo_grad: Callable[[Point], torch.Tensor], geom_steps: int,
                                  learning_rate: float = 0.05,
                                  on_step: Callable[[int, Point], None] | None = None):
    """geometry_opt.py:28-78: recenter, solve, then per step Adam on dE/dR,
    recenter, warm-started solve."""
    nuclei = recenter(initial_nuclei)
    opt = Adam(learning_rate)
    current = solve_at_geometry(nuclei, None)
    trajectory = [current]
    for step in range(geom_steps):
        grad_R = bo_grad(current)
        upd = opt.update(grad_R)
        new_positions = current.nuclei.positions.detach().cpu() + upd
        nuclei = recenter(Nuclei(new_positions, current.nuclei.charges))
        current = solve_at_geometry(nuclei, current)
        if on_step is not None:
            on_step(step, current)
        trajectory.append(current)
    return current, trajectory
