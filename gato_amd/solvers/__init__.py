"""Poisson and Lanczos solvers (reference: src/gato/solvers)."""
from .lanczos import LanczosResult, default_krylov_dim, lanczos
from .poisson import PoissonPlan, hartree_energy, hartree_potential

__all__ = ["LanczosResult", "default_krylov_dim", "lanczos", "PoissonPlan", "hartree_energy",
           "hartree_potential"]
