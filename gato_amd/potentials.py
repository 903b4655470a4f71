"""External potentials sampled on the grid by a HIP kernel (reference:
potentials.py:16-76).  Values are evaluated in double and stored as
``dtype`` (default float64, as in the reference)."""
from __future__ import annotations

import torch

from . import _lib as L
from . import _native as nat
from .geometry import Nuclei
from .grid import Grid3D


def _coulomb_center(R, Z, eps):
    return [L.CENTER_SOFT_COULOMB, float(R[0]), float(R[1]), float(R[2]), float(Z), float(eps),
            0.0, 0.0, 0.0, 0.0]


def softened_coulomb(grid: Grid3D, Z: float = 1.0, epsilon: float | None = None,
                     center=(0.0, 0.0, 0.0), dtype=torch.float64) -> torch.Tensor:
    """-Z / sqrt(|r - r0|^2 + eps^2), eps default h/2 (potentials.py:16-36)."""
    if epsilon is None:
        epsilon = grid.h / 2
    return nat.local_potential(grid.N, grid.L, [_coulomb_center(center, Z, epsilon)], dtype)


def coulomb_centers(nuclei: Nuclei, epsilon: float) -> list:
    P = nuclei.positions.detach().cpu().tolist()
    Z = nuclei.charges.detach().cpu().tolist()
    return [_coulomb_center(P[k], Z[k], epsilon) for k in range(len(Z))]


def multi_center_softened_coulomb(nuclei: Nuclei, grid: Grid3D, epsilon: float | None = None,
                                  dtype=torch.float64) -> torch.Tensor:
    """-sum_k Z_k / sqrt(|r - R_k|^2 + eps^2) (potentials.py:39-61)."""
    if epsilon is None:
        epsilon = grid.h / 2
    return nat.local_potential(grid.N, grid.L, coulomb_centers(nuclei, epsilon), dtype)


def harmonic_oscillator(grid: Grid3D, omega: float = 1.0, dtype=torch.float64) -> torch.Tensor:
    """1/2 omega^2 r^2 (potentials.py:64-71)."""
    c = [L.CENTER_HARMONIC, 0.0, 0.0, 0.0, float(omega), 0.0, 0.0, 0.0, 0.0, 0.0]
    return nat.local_potential(grid.N, grid.L, [c], dtype)


def constant(grid: Grid3D, v0: float = 0.0, dtype=torch.float64) -> torch.Tensor:
    """Uniform potential v0 (potentials.py:74-76)."""
    return torch.full(grid.shape, float(v0), dtype=dtype, device=torch.device("cuda"))
