"""Helium-like atom RHF with a softened nucleus (reference: physics/helium.py:60-144)."""
from __future__ import annotations

from dataclasses import dataclass

import torch

from ..grid import Grid3D
from ..potentials import softened_coulomb
from ..scf import decompose_rhf_energy, scf_rhf


@dataclass
class HeliumResult:
    grid: Grid3D
    orbitals: torch.Tensor
    energy: float
    kinetic: float
    potential: float
    hartree: float
    exchange: float
    virial_ratio: float
    orbital_energies: torch.Tensor
    n_iters: int
    converged: bool
    history: list


def solve_helium(N: int = 64, L: float = 10.0, Z: float = 2.0, epsilon: float | None = None,
                 max_iters: int = 30, tol: float = 1e-6, mixing: float = 0.7, order: int = 4,
                 lanczos_iters: int | None = None, precision: str = "fp64",
                 verbose: bool = False) -> HeliumResult:
    """RHF for one doubly occupied orbital in -Z/sqrt(r^2+eps^2); virial
    2<T>/|<V_ext> + E_H + E_x| as at helium.py:107-112."""
    grid = Grid3D(N=N, L=L)
    V_ext = softened_coulomb(grid, Z=Z, epsilon=epsilon)
    res = scf_rhf(V_ext, grid, 1, max_iters=max_iters, tol=tol, mixing=mixing, order=order,
                  lanczos_iters=lanczos_iters, epsilon=epsilon, precision=precision)
    terms = decompose_rhf_energy(res.orbitals, V_ext, grid, boundary="dirichlet", order=order,
                                 epsilon=epsilon)
    T, V, EH, Ex = (float(terms.kinetic), float(terms.v_ext), float(terms.hartree),
                    float(terms.exchange))
    vr = 2.0 * T / abs(V + EH + Ex)
    if verbose:
        print(f"He Z={Z} N={N} L={L}: E={res.energy:+.6f} iters={res.n_iters} virial={vr:.6f}")
    return HeliumResult(grid, res.orbitals, float(res.energy), T, V, EH, Ex, vr,
                        res.orbital_energies, res.n_iters, res.converged,
                        [(i + 1, e) for i, e in enumerate(res.energy_history)])
