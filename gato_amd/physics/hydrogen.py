"""Hydrogen-like atom (one electron, charge Z at the origin) by imaginary
time (reference: physics/hydrogen.py:36-96).  E0(Z) = -Z^2/2 analytically."""
from __future__ import annotations

from dataclasses import dataclass

import torch

from ..ansatz import init_grid_ansatz
from ..grid import Grid3D, normalize
from ..hamiltonian import Hamiltonian
from ..observables import kinetic_energy, potential_energy, virial_ratio
from ..potentials import softened_coulomb
from ..solvers import imaginary_time


@dataclass
class HydrogenResult:  # hydrogen.py:36-46
    grid: Grid3D
    psi: torch.Tensor
    V: torch.Tensor
    energy: float
    kinetic: float
    potential: float
    virial_ratio: float
    history: list


def solve_hydrogen(N: int = 48, L: float = 10.0, Z: float = 1.0, epsilon: float | None = None,
                   n_steps: int = 3000, order: int = 2, seed: int = 0, verbose: bool = True,
                   dtype: torch.dtype = torch.float64) -> HydrogenResult:
    """Imaginary time from exp(-Z r) (hydrogen.py:49-96); ``seed`` is kept for
    the call signature (the hydrogenic start uses no random numbers)."""
    grid = Grid3D(N=N, L=L)
    V = softened_coulomb(grid, Z=Z, epsilon=epsilon, dtype=dtype)
    H = Hamiltonian(grid=grid, V=V, boundary="dirichlet", order=order)
    psi0 = init_grid_ansatz(grid, init="hydrogenic", Z=Z, dtype=dtype)
    psi, history = imaginary_time(H, psi0, n_steps=n_steps, log_every=max(1, n_steps // 20))
    hist = list(zip(history.steps, history.energies))
    psi = normalize(psi, grid)
    T = kinetic_energy(psi, grid, order=order)
    Vexp = potential_energy(psi, V, grid)
    vr = virial_ratio(psi, V, grid, order=order)
    res = HydrogenResult(grid=grid, psi=psi, V=V, energy=T + Vexp, kinetic=T, potential=Vexp,
                         virial_ratio=vr, history=hist)
    if verbose:
        print(f"Hydrogenic atom Z={Z} on {N}^3 grid, L = {L} (h = {grid.h:.4f}, order {order})")
        print(f"Energy {res.energy:+.6f} Ha (target {-0.5 * Z * Z:+.6f}); virial {vr:+.6f}")
    return res
