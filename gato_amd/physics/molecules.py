"""Diatomic drivers: H2 (all-electron softened Coulomb), LiH and HCl (HGH)
(reference: physics/h2.py:67-136, physics/lih.py:61-189, physics/hcl.py:89-221)."""
from __future__ import annotations

import torch

from ..forces import hellmann_feynman_force as _hf_force
from ..geometry import Nuclei, diatomic_along_z, nuclear_repulsion
from ..grid import Grid3D
from ..potentials import multi_center_softened_coulomb
from ..pseudopotentials import lookup
from ..scf import scf_rhf
from .water import MoleculePoint, solve_rhf_at_geometry


def h2_initial_geometry(R: float = 1.4) -> Nuclei:
    return diatomic_along_z(R, 1.0)


def lih_initial_geometry(R_LiH: float = 3.3) -> Nuclei:
    """Li at -R/2, H at +R/2 (lih.py:61-64)."""
    return diatomic_along_z(R_LiH, 3.0, 1.0)


def lih_hgh_elements():
    return (lookup("Li"), lookup("H"))


def hcl_initial_geometry(R_HCl: float = 2.6) -> Nuclei:
    """H at -R/2, Cl at +R/2 (hcl.py:89-99)."""
    return diatomic_along_z(R_HCl, 1.0, 17.0)


def hcl_hgh_elements():
    return (lookup("H"), lookup("Cl"))


def solve_h2_at_geometry(nuclei: Nuclei, grid: Grid3D, *, max_iters: int = 50, tol: float = 1e-6,
                         mixing: float = 0.7, order: int = 4, lanczos_iters=None,
                         initial_orbitals=None, precision: str = "fp64",
                         **scf_kwargs) -> MoleculePoint:
    """All-electron H2 RHF (h2.py:67-102): V_ext = multi-centre softened Coulomb."""
    V = multi_center_softened_coulomb(nuclei, grid)
    res = scf_rhf(V, grid, 1, initial_orbitals=initial_orbitals, max_iters=max_iters, tol=tol,
                  mixing=mixing, order=order, lanczos_iters=lanczos_iters, precision=precision,
                  **scf_kwargs)
    E_nn = float(nuclear_repulsion(nuclei))
    return MoleculePoint(nuclei, res.orbitals, float(res.energy) + E_nn, float(res.energy), E_nn,
                         res.n_iters, res.converged, res.orbital_energies, res.energy_history)


def solve_lih_at_geometry(nuclei, grid, **kw) -> MoleculePoint:
    return solve_rhf_at_geometry(nuclei, lih_hgh_elements(), grid, **kw)


def solve_hcl_at_geometry(nuclei, grid, **kw) -> MoleculePoint:
    return solve_rhf_at_geometry(nuclei, hcl_hgh_elements(), grid, **kw)


def h2_force(point: MoleculePoint, grid: Grid3D, order: int = 4) -> torch.Tensor:
    """Hellmann-Feynman force of all-electron H2 (h2.py:109-136)."""
    return _hf_force(point.nuclei.positions, point.nuclei.charges, point.orbitals, None, grid, order)


def optimize_diatomic_geometry(initial_nuclei: Nuclei, elements, grid: Grid3D, *,
                               geom_steps: int = 30, learning_rate: float = 0.05,
                               scf_max_iters: int = 60, scf_tol: float = 1e-6,
                               scf_mixing: float = 0.5, order: int = 4,
                               lanczos_iters: int | None = None, verbose: bool = True,
                               precision: str = "fp64"):
    """Alternating SCF + Adam on Hellmann-Feynman gradients for an HGH diatomic
    (hcl.py:234-290, lih.py: same pattern); trajectory of (step, E, R)."""
    from ..geometry import bond_length
    from ..geometry_opt import optimize_geometry_alternating
    from .water import WaterGeomOptResult, hellmann_feynman_force

    def solve_at(nuclei, prev):
        init = prev.orbitals if prev is not None else None
        return solve_rhf_at_geometry(nuclei, elements, grid, initial_orbitals=init,
                                     max_iters=scf_max_iters, tol=scf_tol, mixing=scf_mixing,
                                     order=order, lanczos_iters=lanczos_iters, precision=precision)

    def bo_grad(p):
        return -hellmann_feynman_force(p, elements, grid, order)

    trajectory = []

    def on_step(step, p):
        R = float(bond_length(p.nuclei, 0, 1))
        trajectory.append((step, p.energy, R))
        if verbose:
            print(f"  geom step {step:3d}  E = {p.energy:+.6f}  R = {R:.4f}", flush=True)

    final, _ = optimize_geometry_alternating(initial_nuclei, solve_at_geometry=solve_at,
                                             bo_grad=bo_grad, geom_steps=geom_steps,
                                             learning_rate=learning_rate, on_step=on_step)
    return WaterGeomOptResult(final, trajectory)
