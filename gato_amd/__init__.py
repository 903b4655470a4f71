"""gato_amd: MI355X-native real-space RHF hot path (Hamiltonian apply,
Hockney Poisson for Hartree and exchange, Lanczos, SCF, energy and
Hellmann-Feynman forces) behind the call surface of the reference package
``gato`` (src/gato/__init__.py:6-56).

Fields live on the GPU as torch tensors; every operation runs hand-written
HIP kernels of ``libgato_amd.so`` through its C ABI (include/gato_amd.h).
"""
from . import _lib
from .geometry import (Nuclei, bond_angle, bond_length, center_of_mass, diatomic_along_z,
                       make_nuclei, nuclear_repulsion, recenter)
from .grid import Grid3D, inner_product, integrate, norm_sq, normalize
from .hamiltonian import Hamiltonian
from .operators import kinetic, laplacian
from .potentials import (constant, harmonic_oscillator, multi_center_softened_coulomb,
                         softened_coulomb)
from .pseudopotentials import (HGHParams, HGHProjector, NonlocalPP, hgh_apply,
                               hgh_local_potential, hgh_nonlocal_apply, lookup,
                               total_valence_charge)
from .scf import (RHFEnergyTerms, RHFFock, SCFResult, decompose_rhf_energy, exchange_apply,
                  rhf_energy, scf_rhf)
from .solvers import (ImagTimeHistory, LanczosResult, PoissonPlan, block_lanczos,
                      default_krylov_dim, hartree_energy, hartree_potential, imaginary_time,
                      lanczos)
from .ansatz import init_grid_ansatz, init_lcao
from .forces import bo_energy, hellmann_feynman_force
from .geometry_opt import Adam, optimize_geometry_alternating
from .orbital_grad import rayleigh_descent, rayleigh_quotient, rhf_energy_ad

__version__ = "0.1.0"


def load_native():
    """Load the HIP library (raises if it was not built)."""
    return _lib.load()


def set_autotune(on: bool = True) -> None:
    """Switch the stencil's per-shape kernel timing on (the default) or off.
    Off, every shape takes the measured static rule; results are the same
    either way (every variant evaluates the same expression in the same
    order)."""
    _lib.lib().gato_set_autotune(int(bool(on)))
