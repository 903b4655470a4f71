"""Free-space Hartree potential by the Hockney doubled-grid method
(reference: solvers/poisson.py:1-92).

J(r) = int rho(r') / sqrt(|r-r'|^2 + eps^2) dV', eps default h/2, evaluated
as a circular convolution on the (2N)^3 grid with the softened kernel of
poisson.py:29-44.  The transform of the kernel is built once per
(N, h, eps, dtype) on the device and cached in a :class:`PoissonPlan`; each
solve is five batched, pruned LDS-staged FFT passes (two real solves per
complex transform), see gato_amd/csrc/poisson.hip.
"""
from __future__ import annotations

import ctypes as C
import threading

import torch

from .. import _lib as L
from .. import _native as nat
from ..grid import Grid3D


class PoissonPlan:
    """Device-resident plan for one grid, softening and dtype."""

    def __init__(self, N: int, h: float, epsilon: float, dtype: torch.dtype = torch.float64,
                 device=None, max_pairs: int | None = None):
        self.N, self.h, self.epsilon, self.dtype = int(N), float(h), float(epsilon), dtype
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if max_pairs is None:
            max_pairs = 8 if N <= 160 else 2
        self.max_pairs = int(max_pairs)
        self._h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.check(L.lib().gato_poisson_plan_create(
                L.dtype_code(dtype), self.N, self.h, self.epsilon, self.max_pairs,
                C.byref(self._h), L.stream_ptr()))

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def nbytes(self) -> int:
        """Device memory held: the two pitched pair workspaces (row pitch
        2N + 16), the folded kernel transform and the twiddles."""
        N, M = self.N, 2 * self.N
        r = torch.finfo(self.dtype).bits // 8
        work = self.max_pairs * N * (N + M) * (M + 16) * 2 * r
        return work + (M // 2 + 1) ** 3 * r + M * 2 * r

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            L.lib().gato_poisson_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fields(self, x: torch.Tensor, name: str) -> torch.Tensor:
        x = nat.contig(x, name)
        if x.dtype != self.dtype:
            raise ValueError(f"{name} dtype {x.dtype} does not match the plan ({self.dtype})")
        if tuple(x.shape[-3:]) != (self.N,) * 3:
            raise ValueError(f"{name} shape {tuple(x.shape)} does not match N={self.N}")
        return x

    def hartree_potential(self, rho: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        rho = self._fields(rho, "rho")
        n3 = self.N ** 3
        k = rho.numel() // n3
        if out is None:
            out = torch.empty_like(rho)
        L.check(L.lib().gato_hartree_potential(self._h, k, rho.data_ptr(), n3, out.data_ptr(), n3,
                                               L.stream_ptr()))
        return out

    def hartree_energy(self, rho: torch.Tensor) -> torch.Tensor:
        """0.5 int rho J[rho] for each density of a (k, N, N, N) stack -> (k,) float64."""
        rho = self._fields(rho, "rho")
        n3 = self.N ** 3
        k = rho.numel() // n3
        E = torch.empty(k, dtype=torch.float64, device=rho.device)
        ws = L.workspace(rho.device)
        L.check(L.lib().gato_hartree_energy(self._h, k, rho.data_ptr(), n3, E.data_ptr(),
                                            ws.data_ptr(), L.stream_ptr()))
        return E

    def exchange_apply(self, psi: torch.Tensor, orbitals_knnn: torch.Tensor, coef: float = 1.0,
                       out: torch.Tensor | None = None, accumulate: bool = False) -> torch.Tensor:
        """out (+)= coef * sum_j phi_j J[phi_j psi]; orbitals as (n, N, N, N)."""
        psi = self._fields(psi, "psi")
        orb = self._fields(orbitals_knnn, "orbitals")
        n3 = self.N ** 3
        if out is None:
            out = torch.empty_like(psi)
            accumulate = False
        L.check(L.lib().gato_exchange_apply(self._h, psi.data_ptr(), orb.data_ptr(), n3,
                                            orb.numel() // n3, float(coef), int(accumulate),
                                            out.data_ptr(), L.stream_ptr()))
        return out

    def pair_integrals(self, orbitals_knnn: torch.Tensor) -> torch.Tensor:
        """(ij|ij) for i <= j in row-major upper-triangle order, float64."""
        orb = self._fields(orbitals_knnn, "orbitals")
        n3 = self.N ** 3
        n = orb.numel() // n3
        out = torch.empty(n * (n + 1) // 2, dtype=torch.float64, device=orb.device)
        ws = L.workspace(orb.device)
        L.check(L.lib().gato_exchange_pair_integrals(self._h, orb.data_ptr(), n3, n,
                                                     out.data_ptr(), ws.data_ptr(),
                                                     L.stream_ptr()))
        return out


_PLANS: dict = {}
_PLANS_LOCK = threading.Lock()  # host threads of a batch share the cache
_PLAN_CACHE_BYTES = 32 << 30  # of the GPU's 288 GB


def get_plan(grid: Grid3D, epsilon: float | None, dtype: torch.dtype, device=None) -> PoissonPlan:
    """The cached plan for this grid, softening and dtype on the current
    stream.  A plan's FFT scratch serves one stream at a time, so every stream
    has its own plans.  The cache holds up to 32 GiB of plans, oldest out
    first; a plan dropped from it is freed when its last user releases it
    (never while a solve still holds it), or at the end of a
    ``deferred_release()`` scope."""
    if epsilon is None:
        epsilon = grid.h / 2
    idx = L.device_index(device)
    dev = torch.device("cuda", idx)
    key = (grid.N, float(grid.h), float(epsilon), dtype, idx, L.stream_ptr(idx))
    with _PLANS_LOCK:
        p = _PLANS.get(key)
    if p is None:
        p = PoissonPlan(grid.N, grid.h, epsilon, dtype, dev)  # outside the lock: slow
        with _PLANS_LOCK:
            while _PLANS and sum(q.nbytes for q in _PLANS.values()) + p.nbytes > _PLAN_CACHE_BYTES:
                L.retire(_PLANS.pop(next(iter(_PLANS))))
            _PLANS[key] = p
    return p


def hartree_potential(rho: torch.Tensor, grid: Grid3D, epsilon: float | None = None) -> torch.Tensor:
    """J = int rho(r') G(r - r') dV' (poisson.py:60-79)."""
    return get_plan(grid, epsilon, rho.dtype, rho.device).hartree_potential(rho)


def hartree_energy(rho: torch.Tensor, grid: Grid3D, epsilon: float | None = None) -> torch.Tensor:
    """E_H = 1/2 int rho J (poisson.py:82-92); 0-dim float64 device tensor."""
    return get_plan(grid, epsilon, rho.dtype, rho.device).hartree_energy(rho)[0]
