"""H2O with HGH pseudopotentials: RHF at a geometry, BO energy, Hellmann-Feynman
force and geometry optimisation (reference: physics/water.py:66-318)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from ..forces import bo_energy as _bo_energy
from ..forces import hellmann_feynman_force as _hf_force
from ..geometry import Nuclei, bond_angle, bond_length, make_nuclei, nuclear_repulsion
from ..geometry_opt import optimize_geometry_alternating
from ..grid import Grid3D
from ..pseudopotentials import (HGHParams, hgh_local_potential, lookup, nonlocal_plan,
                                total_valence_charge)
from ..scf import scf_rhf


def water_initial_geometry(R_OH: float = 1.85, angle_deg: float = 100.0) -> Nuclei:
    """O at the origin, H at (+-R sin(a/2), 0, R cos(a/2)) (water.py:66-82)."""
    half = math.radians(angle_deg) / 2.0
    return make_nuclei([[0.0, 0.0, 0.0],
                        [R_OH * math.sin(half), 0.0, R_OH * math.cos(half)],
                        [-R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]], [8.0, 1.0, 1.0])


def water_hgh_elements() -> tuple:
    return (lookup("O"), lookup("H"), lookup("H"))


def pp_nuclear_repulsion(nuclei: Nuclei, elements) -> torch.Tensor:
    """E_nn on the valence charges Z_ion (water.py:100-110)."""
    if nuclei.n < 2:
        return torch.zeros((), dtype=torch.float64)
    zion = torch.tensor([p.Z_ion for p in elements], dtype=torch.float64)
    return nuclear_repulsion(Nuclei(nuclei.positions, zion))


@dataclass
class MoleculePoint:
    nuclei: Nuclei
    orbitals: torch.Tensor
    energy: float
    rhf_energy: float
    nuclear_repulsion: float
    n_iters: int
    converged: bool
    orbital_energies: torch.Tensor
    energy_history: list = None


WaterPoint = MoleculePoint


def solve_rhf_at_geometry(nuclei: Nuclei, elements, grid: Grid3D, *, max_iters: int = 60,
                          tol: float = 1e-6, mixing: float = 0.5, order: int = 4,
                          lanczos_iters: int | None = None, initial_orbitals=None,
                          precision: str = "fp64", **scf_kwargs) -> MoleculePoint:
    """RHF with HGH PPs at fixed nuclei (water.py:129-174)."""
    n_el = int(total_valence_charge(elements))
    if n_el % 2 != 0:
        raise ValueError(f"Closed-shell RHF requires even electron count; got {n_el}")
    V_local = hgh_local_potential(nuclei, elements, grid)
    V_nl = nonlocal_plan(nuclei, elements, grid, torch.float64) if any(p.projectors for p in elements) else None
    res = scf_rhf(V_local, grid, n_el // 2, initial_orbitals=initial_orbitals, max_iters=max_iters,
                  tol=tol, mixing=mixing, order=order, lanczos_iters=lanczos_iters,
                  V_nl_apply=V_nl, precision=precision, **scf_kwargs)
    E_nn = float(pp_nuclear_repulsion(nuclei, elements))
    return MoleculePoint(nuclei, res.orbitals, float(res.energy) + E_nn, float(res.energy), E_nn,
                         res.n_iters, res.converged, res.orbital_energies, res.energy_history)


def bo_energy(positions, charges, orbitals, elements, grid: Grid3D, order: int = 4):
    """E(R) at fixed orbitals (water.py:181-215); differentiable in positions."""
    return _bo_energy(positions, charges, orbitals, elements, grid, order)


def hellmann_feynman_force(point: MoleculePoint, elements, grid: Grid3D, order: int = 4):
    """(K, 3) force at converged orbitals (water.py:227-241)."""
    return _hf_force(point.nuclei.positions, point.nuclei.charges, point.orbitals, elements, grid,
                     order)


@dataclass
class WaterGeomOptResult:
    final: MoleculePoint
    trajectory: list


def optimize_water_geometry(initial_nuclei: Nuclei, elements, grid: Grid3D, *, geom_steps: int = 30,
                            learning_rate: float = 0.05, scf_max_iters: int = 60,
                            scf_tol: float = 1e-6, scf_mixing: float = 0.5, order: int = 4,
                            lanczos_iters: int | None = None, verbose: bool = True,
                            precision: str = "fp64") -> WaterGeomOptResult:
    """Alternating SCF + Adam on Hellmann-Feynman gradients (water.py:254-318)."""

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
        ang = math.degrees(float(bond_angle(p.nuclei, 1, 0, 2)))
        trajectory.append((step, p.energy, R, ang))
        if verbose:
            print(f"  geom step {step:3d}  E = {p.energy:+.6f}  R_OH = {R:.4f}  angle = {ang:.2f}  "
                  f"SCF iters = {p.n_iters}", flush=True)

    final, _ = optimize_geometry_alternating(initial_nuclei, solve_at_geometry=solve_at,
                                             bo_grad=bo_grad, geom_steps=geom_steps,
                                             learning_rate=learning_rate, on_step=on_step)
    return WaterGeomOptResult(final, trajectory)
