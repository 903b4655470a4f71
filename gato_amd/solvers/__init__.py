"""Poisson, Lanczos and imaginary-time solvers (reference: src/gato/solvers)."""
from .block_lanczos import block_lanczos
from .imag_time import ImagTimeHistory, imaginary_time
from .lanczos import LanczosResult, default_krylov_dim, lanczos
from .poisson import PoissonPlan, hartree_energy, hartree_potential

__all__ = ["block_lanczos", "ImagTimeHistory", "imaginary_time", "LanczosResult", "default_krylov_dim", "lanczos",
           "PoissonPlan", "hartree_energy", "hartree_potential"]
