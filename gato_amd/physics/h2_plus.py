"""H2+ (one electron, two protons): Born-Oppenheimer energy, Hellmann-Feynman
force, BO curve and geometry optimisation with the imaginary-time solver
(reference: physics/h2_plus.py:66-228).

The electronic energy <psi|T + V(R)|psi>/<psi|psi> and its R-derivative run
on the GPU: the stencil for T, the native softened-Coulomb field for V(R),
and the hand-written adjoint (local-potential gradient against the density
|psi|^2 / <psi|psi>) behind ``torch.autograd`` for the force.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from .. import _native as nat
from ..ansatz import init_lcao
from ..forces import _ExternalEnergy
from ..geometry import Nuclei, bond_length, diatomic_along_z, nuclear_repulsion
from ..geometry_opt import optimize_geometry_alternating
from ..grid import Grid3D, inner_product, norm_sq, normalize
from ..hamiltonian import Hamiltonian
from ..operators import kinetic
from ..potentials import multi_center_softened_coulomb
from ..solvers import imaginary_time


def bo_energy(positions: torch.Tensor, charges: torch.Tensor, psi: torch.Tensor, grid: Grid3D,
              epsilon: float, order: int = 4) -> torch.Tensor:
    """<psi|T + V(R)|psi>/<psi|psi> + E_nn(R), differentiable in ``positions``
    (h2_plus.py:66-88).  A float64 scalar on the host."""
    P = positions.to(torch.float64)
    Zc = charges.to(torch.float64)
    psi = nat.contig(psi, "psi")
    n2 = float(norm_sq(psi, grid))
    e_kin = float(inner_product(psi, kinetic(psi, grid.h, "dirichlet", order), grid)) / n2
    stack = psi.reshape((1,) + tuple(psi.shape))
    rho2 = nat.density_mix(stack, 1.0, None)  # 2 |psi|^2
    e_ext = (0.5 / n2) * _ExternalEnergy.apply(P, Zc, stack, rho2, grid, None, epsilon)
    return e_kin + e_ext + nuclear_repulsion(Nuclei(P, Zc))


def hellmann_feynman_force(positions: torch.Tensor, charges: torch.Tensor, psi: torch.Tensor,
                           grid: Grid3D, epsilon: float, order: int = 4) -> torch.Tensor:
    """F = -dE/dR at fixed psi, shape (K, 3) (h2_plus.py:99-108)."""
    P = positions.detach().to(torch.float64).clone().requires_grad_(True)
    E = bo_energy(P, charges, psi, grid, epsilon, order)
    (g,) = torch.autograd.grad(E, P)
    return -g


@dataclass
class SolvePoint:  # h2_plus.py:115-121
    nuclei: Nuclei
    psi: torch.Tensor
    energy: float
    electronic_energy: float
    nuclear_repulsion: float


def solve_electronic(nuclei: Nuclei, grid: Grid3D, epsilon: float | None = None,
                     n_steps: int = 2000, order: int = 4,
                     dtype: torch.dtype = torch.float64) -> SolvePoint:
    """Ground state at fixed nuclei by imaginary time from the LCAO guess
    (h2_plus.py:124-146)."""
    if epsilon is None:
        epsilon = grid.h / 2
    V = multi_center_softened_coulomb(nuclei, grid, epsilon, dtype=dtype)
    H = Hamiltonian(grid=grid, V=V, boundary="dirichlet", order=order)
    psi0 = normalize(init_lcao(nuclei, grid, dtype=dtype), grid)
    psi, _ = imaginary_time(H, psi0, n_steps=n_steps, log_every=n_steps + 1)
    E_nn = float(nuclear_repulsion(nuclei))
    E_el = float(H.rayleigh(psi))
    return SolvePoint(nuclei=nuclei, psi=psi, energy=E_el + E_nn, electronic_energy=E_el,
                      nuclear_repulsion=E_nn)


def bo_curve(R_values, grid: Grid3D, Z: float = 1.0, epsilon: float | None = None,
             n_steps: int = 2000, order: int = 4) -> list:
    """E(R) of a homonuclear diatomic along z, atoms at +-R/2 (h2_plus.py:153-168)."""
    return [solve_electronic(diatomic_along_z(float(R), Z1=Z), grid, epsilon=epsilon,
                             n_steps=n_steps, order=order) for R in R_values]


@dataclass
class GeomOptResult:  # h2_plus.py:175-178
    final: SolvePoint
    trajectory: list


def optimize_geometry(initial_nuclei: Nuclei, grid: Grid3D, epsilon: float | None = None,
                      electronic_steps: int = 2000, geom_steps: int = 40,
                      learning_rate: float = 0.05, order: int = 4,
                      verbose: bool = False) -> GeomOptResult:
    """Alternating relaxation: converge psi, one Adam step on R, recenter
    (h2_plus.py:181-228)."""
    if epsilon is None:
        epsilon = grid.h / 2

    def solve_at(nuclei: Nuclei, _prev) -> SolvePoint:
        return solve_electronic(nuclei, grid, epsilon=epsilon, n_steps=electronic_steps,
                                order=order)

    def bo_grad(point: SolvePoint) -> torch.Tensor:
        return -hellmann_feynman_force(point.nuclei.positions, point.nuclei.charges, point.psi,
                                       grid, epsilon, order)

    trajectory: list = []

    def on_step(step: int, p: SolvePoint) -> None:
        R_ij = float(bond_length(p.nuclei, 0, 1)) if p.nuclei.n >= 2 else float("nan")
        trajectory.append((step, p.energy, R_ij))
        if verbose:
            print(f"  geom step {step:3d}  E = {p.energy:+.6f}  R01 = {R_ij:.4f}")

    final, _ = optimize_geometry_alternating(initial_nuclei, solve_at_geometry=solve_at,
                                             bo_grad=bo_grad, geom_steps=geom_steps,
                                             learning_rate=learning_rate, on_step=on_step)
    return GeomOptResult(final=final, trajectory=trajectory)
