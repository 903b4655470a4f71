"""Born-Oppenheimer energy at fixed orbitals and its Hellmann-Feynman force
(reference: physics/water.py:181-241 and the same pattern in lih.py, hcl.py,
h2.py).

Only V_loc(R), V_nl(R) and E_nn(R) depend on the nuclear positions.  Their
derivative is a hand-written adjoint on the GPU behind a
``torch.autograd.Function``: the backward pass recomputes dV_loc/dR and the
projector derivatives point by point (nothing field-sized is stored) and
reduces them against the density / the orbitals in double.  E_nn is a
K x K host expression differentiated by autograd.
"""
from __future__ import annotations

import torch

from . import _native as nat
from .geometry import Nuclei, nuclear_repulsion
from .grid import Grid3D
from .potentials import coulomb_centers
from .pseudopotentials import NonlocalPP, local_centers
from .scf import _energy_terms, to_stack


def _field_centers(positions: torch.Tensor, charges: torch.Tensor, elements, epsilon):
    nuc = Nuclei(positions.detach().cpu(), charges.detach().cpu())
    if elements is None:
        return coulomb_centers(nuc, epsilon)
    return local_centers(nuc, elements)


class _ExternalEnergy(torch.autograd.Function):
    """E_ext(R) = int rho V_loc(R) + 2 sum_i <phi_i|V_nl(R)|phi_i>  (float64 scalar)."""

    @staticmethod
    def forward(ctx, positions, charges, orb, rho, grid, elements, epsilon):
        centers = _field_centers(positions, charges, elements, epsilon)
        V = nat.local_potential(grid.N, grid.L, centers, orb.dtype, device=orb.device)
        n3 = grid.N ** 3
        e_loc =nat.dot_batched(V, rho, 1, 0, 0, scale=grid.dV, n=n3)[0]
        e_nl = torch.zeros((), dtype=torch.float64, device=orb.device)
        nl = None
        if elements is not None and any(p.projectors for p in elements):
            nl = NonlocalPP(Nuclei(positions.detach().cpu(), charges.detach().cpu()), elements,
                            grid, orb.dtype)
            e_nl = 2.0 * nl.expectation(orb).sum()
        ctx.save_for_backward(orb, rho)
        ctx.grid, ctx.centers, ctx.nl = grid, centers, nl
        return (e_loc + e_nl).cpu()

    @staticmethod
    def backward(ctx, g):
        orb, rho = ctx.saved_tensors
        grid = ctx.grid
        dE = nat.local_potential_grad(grid.N, grid.L, ctx.centers, rho)
        if ctx.nl is not None:
            dE = dE + ctx.nl.grad(orb, occ=2.0)
        return (g * dE.cpu()), None, None, None, None, None, None


def bo_energy(positions: torch.Tensor, charges: torch.Tensor, orbitals: torch.Tensor, elements,
              grid: Grid3D, order: int = 4, epsilon: float | None = None) -> torch.Tensor:
    """Total BO energy at fixed orbitals, differentiable in ``positions``
    (water.py:181-215).  ``elements`` None means all-electron softened Coulomb
    nuclei (h2.py / helium.py); otherwise HGH with E_nn on Z_ion."""
    orb = to_stack(orbitals)
    eps_ext = grid.h / 2
    P = positions.to(torch.float64)
    Zc = charges.to(torch.float64)
    # R-independent part: T + E_H + E_x (the R-dependent terms come from the adjoint)
    zero_V = torch.zeros(grid.shape, dtype=orb.dtype, device=orb.device)
    terms = _energy_terms(orb, zero_V, grid, "dirichlet", order, epsilon, None)
    e_int = torch.tensor(terms.kinetic + terms.hartree + terms.exchange, dtype=torch.float64)
    rho = nat.density_mix(orb, 1.0, None)
    e_ext = _ExternalEnergy.apply(P, Zc, orb, rho, grid, elements, eps_ext)
    if elements is None:
        e_nn = nuclear_repulsion(Nuclei(P, Zc))
    else:
        zion = torch.tensor([p.Z_ion for p in elements], dtype=torch.float64)
        e_nn = nuclear_repulsion(Nuclei(P, zion)) if len(elements) >= 2 else torch.zeros((), dtype=torch.float64)
    return e_int + e_ext + e_nn


def hellmann_feynman_force(positions: torch.Tensor, charges: torch.Tensor, orbitals: torch.Tensor,
                           elements, grid: Grid3D, order: int = 4) -> torch.Tensor:
    """F = -dE/dR at fixed orbitals (water.py:227-241), shape (K, 3)."""
    P = positions.detach().to(torch.float64).clone().requires_grad_(True)
    E = bo_energy(P, charges, orbitals, elements, grid, order)
    (g,) = torch.autograd.grad(E, P)
    return -g
