"""Single-particle Hamiltonian H = T + V as a matrix-free operator
(reference: hamiltonian.py:18-35).  ``apply`` is one fused stencil launch;
``lanczos`` drives it natively through ``_native_fock``."""
from __future__ import annotations

import contextlib
import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib as L
from . import _native as nat
from .grid import Grid3D, inner_product, norm_sq


@dataclass(frozen=True)
class Hamiltonian:
    grid: Grid3D
    V: torch.Tensor
    boundary: str = "dirichlet"
    order: int = 2  # hamiltonian.py:23

    def _V_as(self, dtype: torch.dtype) -> torch.Tensor:
        V = self.V
        if V.dtype != dtype:
            V = nat.convert(V, dtype)
        return nat.contig(V, "V")

    def apply(self, psi: torch.Tensor) -> torch.Tensor:
        """H psi = (-1/2 lap + V) psi (hamiltonian.py:25-27)."""
        return nat.stencil(psi, self.grid.h, self.order, self.boundary, -0.5,
                           V=self._V_as(psi.dtype))

    def expectation(self, psi: torch.Tensor) -> torch.Tensor:
        return inner_product(psi, self.apply(psi), self.grid)

    def rayleigh(self, psi: torch.Tensor) -> torch.Tensor:
        return self.expectation(psi) / norm_sq(psi, self.grid)

    @property
    def _native_fock(self):
        @contextlib.contextmanager
        def ctx(dtype):
            V = self._V_as(dtype)
            h = C.c_void_p()
            L.check(L.lib().gato_fock_create(
                L.dtype_code(dtype), self.grid.N, self.grid.h, nat.check_order(self.order),
                nat.boundary_code(self.boundary), V.data_ptr(), None, None, 0, 0, None,
                C.byref(h)))
            try:
                yield h
            finally:
                L.lib().gato_fock_destroy(h)
                del V
        return ctx
