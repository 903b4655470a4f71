"""Cell-centred 3D grid and midpoint-rule integrals (reference: grid.py:1-73).

Grid points sit at x_i = -L/2 + (i + 1/2) h, i = 0..N-1, h = L/N; fields are
(N, N, N) GPU tensors indexed [x][y][z] with z contiguous.  Integrals and inner
products run on the device in double with compensated summation and return
0-dim float64 device tensors (``float(x)`` syncs, like the reference's
``float(jax_array)``).
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _native as nat


def default_device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


@dataclass(frozen=True)
class Grid3D:
    """Uniform cell-centred grid on [-L/2, L/2]^3 (grid.py:23-54)."""

    N: int
    L: float

    @property
    def h(self) -> float:
        return self.L / self.N

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.N, self.N, self.N)

    @property
    def dV(self) -> float:
        return self.h ** 3

    def axis(self, dtype=torch.float64, device=None) -> torch.Tensor:
        dev = default_device() if device is None else device
        return -self.L / 2 + (torch.arange(self.N, dtype=torch.float64, device=dev) + 0.5) * self.h

    def coords(self, dtype=torch.float64, device=None):
        a = self.axis(device=device)
        X, Y, Z = torch.meshgrid(a, a, a, indexing="ij")
        return X.to(dtype), Y.to(dtype), Z.to(dtype)

    def radial(self, softening: float = 0.0, dtype=torch.float64, device=None) -> torch.Tensor:
        X, Y, Z = self.coords(device=device)
        return torch.sqrt(X * X + Y * Y + Z * Z + softening * softening).to(dtype)


def integrate(psi: torch.Tensor, grid: Grid3D) -> torch.Tensor:
    """int psi dV (grid.py:57-59)."""
    ones = torch.ones_like(psi)
    return nat.dot_batched(ones, psi, 1, 0, 0, scale=grid.dV, n=psi.numel())[0]


def inner_product(phi: torch.Tensor, psi: torch.Tensor, grid: Grid3D) -> torch.Tensor:
    """<phi|psi> = int phi psi dV for real fields (grid.py:62-64)."""
    if phi.shape != psi.shape:
        raise ValueError("inner_product operands must have the same shape")
    return nat.dot_batched(phi, psi, 1, 0, 0, scale=grid.dV, n=psi.numel())[0]


def norm_sq(psi: torch.Tensor, grid: Grid3D) -> torch.Tensor:
    """<psi|psi> (grid.py:67-69)."""
    return nat.dot_batched(psi, psi, 1, 0, 0, scale=grid.dV, n=psi.numel())[0]


def normalize(psi: torch.Tensor, grid: Grid3D) -> torch.Tensor:
    """psi / ||psi|| (grid.py:72-73), as a new tensor."""
    return nat.normalize(psi, grid.dV)
