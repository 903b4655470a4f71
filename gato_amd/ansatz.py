"""Initial grid wave functions (reference: ansatz/grid.py:15-76).

The analytic guesses are evaluated by the native local-field kernel (one
hydrogenic-1s or harmonic centre per nucleus); only the random start uses a
torch generator, since the reference's jax.random stream cannot be
reproduced outside JAX anyway.
"""
from __future__ import annotations

import torch

from . import _lib as L
from . import _native as nat
from .geometry import Nuclei
from .grid import Grid3D


def _slater_center(R, Z: float, softening: float) -> list:
    return [L.CENTER_SLATER_1S, float(R[0]), float(R[1]), float(R[2]), float(Z), float(softening),
            0.0, 0.0, 0.0, 0.0]


def init_lcao(nuclei: Nuclei, grid: Grid3D, softening: float | None = None,
              dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """psi0(r) = sum_k Z_k^{3/2} exp(-Z_k sqrt(|r - R_k|^2 + s^2)), unnormalised
    (ansatz/grid.py:53-76); s defaults to h/2."""
    if softening is None:
        softening = grid.h / 2
    P = nuclei.positions.detach().cpu().to(torch.float64)
    Z = nuclei.charges.detach().cpu().to(torch.float64)
    centers = [_slater_center(P[k], float(Z[k]), softening) for k in range(P.shape[0])]
    return nat.local_potential(grid.N, grid.L, centers, dtype)


def init_grid_ansatz(grid: Grid3D, generator: torch.Generator | None = None,
                     init: str = "hydrogenic", Z: float = 1.0, noise: float = 0.0,
                     dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """ansatz/grid.py:15-46: "hydrogenic" exp(-Z r) with r softened by h/2,
    "gaussian" exp(-r^2/2), "random" unit normal noise; ``noise`` adds that
    much Gaussian noise to the analytic guesses."""
    dev = torch.device("cuda")
    if init == "hydrogenic":
        c = _slater_center((0.0, 0.0, 0.0), Z, grid.h / 2)
        psi = nat.local_potential(grid.N, grid.L, [c], dtype)
        psi = nat.axpby(Z ** -1.5, psi, 0.0, psi)  # drop the Z^{3/2} prefactor
    elif init == "gaussian":
        c = [L.CENTER_HARMONIC, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]
        psi = torch.exp(-nat.local_potential(grid.N, grid.L, [c], dtype))  # exp(-r^2/2)
    elif init == "random":
        return torch.randn(grid.shape, generator=generator, dtype=dtype, device=dev)
    else:
        raise ValueError(f"Unknown init {init!r}")
    if noise > 0:
        psi = psi + noise * torch.randn(grid.shape, generator=generator, dtype=dtype, device=dev)
    return psi
