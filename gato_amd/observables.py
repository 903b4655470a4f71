"""Observables: kinetic / potential energy, virial ratio, radial density
(reference: observables.py:15-64).  All divide by <psi|psi>."""
from __future__ import annotations

import torch

from . import _native as nat
from .grid import Grid3D, inner_product, norm_sq
from .operators import kinetic


def kinetic_energy(psi: torch.Tensor, grid: Grid3D, boundary: str = "dirichlet",
                   order: int = 2) -> float:
    """<T> = <psi|-1/2 lap|psi> / <psi|psi> (observables.py:15-21)."""
    return float(inner_product(psi, kinetic(psi, grid.h, boundary, order), grid)) / \
        float(norm_sq(psi, grid))


def potential_energy(psi: torch.Tensor, V: torch.Tensor, grid: Grid3D) -> float:
    """<V> = <psi|V|psi> / <psi|psi> (observables.py:24-27): one fused
    stencil-free multiply (c_lap = 0) and a dot."""
    Vpsi = nat.stencil(psi, grid.h, 2, "dirichlet", 0.0, V=nat.contig(V.to(psi.dtype), "V"))
    return float(inner_product(psi, Vpsi, grid)) / float(norm_sq(psi, grid))


def virial_ratio(psi: torch.Tensor, V: torch.Tensor, grid: Grid3D, boundary: str = "dirichlet",
                 order: int = 2) -> float:
    """2<T> / |<V>| (observables.py:30-42); 1 for an exact Coulomb eigenstate."""
    return 2.0 * kinetic_energy(psi, grid, boundary, order) / abs(potential_energy(psi, V, grid))


def radial_density(psi: torch.Tensor, grid: Grid3D, n_bins: int = 100,
                   r_max: float | None = None):
    """Radial probability density P(r), integral = <psi|psi> (observables.py:45-64).
    Analysis helper (histogram of |psi|^2 dV over r), off the hot path."""
    r = grid.radial(device=psi.device).reshape(-1)
    dens = (psi.double() ** 2 * grid.dV).reshape(-1)
    if r_max is None:
        r_max = grid.L / 2
    bins = torch.linspace(0.0, r_max, n_bins + 1, dtype=torch.float64, device=psi.device)
    idx = torch.bucketize(r, bins, right=True) - 1
    keep = (idx >= 0) & (idx < n_bins)
    hist = torch.zeros(n_bins, dtype=torch.float64, device=psi.device)
    hist.index_add_(0, idx[keep], dens[keep])
    centers = 0.5 * (bins[1:] + bins[:-1])
    return centers.cpu(), (hist / (bins[1:] - bins[:-1])).cpu()
