"""Matrix-free finite-difference Laplacian and kinetic operator
(reference: operators.py:40-130), on the fused LDS-tiled HIP stencil.

order 2: 7-point (1, -6)/h^2; order 4: 13-point (-1/12, 16/12, -30/12 per
axis)/h^2.  boundary "dirichlet" treats psi as zero outside the grid,
"periodic" wraps.  psi may carry leading batch dimensions (..., N, N, N).
"""
from __future__ import annotations

import torch

from . import _native as nat


def laplacian(psi: torch.Tensor, h: float, boundary: str = "dirichlet", order: int = 2) -> torch.Tensor:
    """Central-difference 3D Laplacian (operators.py:40-64)."""
    nat.check_order(order)
    nat.boundary_code(boundary)
    return nat.stencil(psi, h, order, boundary, 1.0)


def kinetic(psi: torch.Tensor, h: float, boundary: str = "dirichlet", order: int = 2) -> torch.Tensor:
    """T psi = -1/2 lap psi (operators.py:122-130)."""
    nat.check_order(order)
    nat.boundary_code(boundary)
    return nat.stencil(psi, h, order, boundary, -0.5)
