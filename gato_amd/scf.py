"""Closed-shell RHF on the grid (reference: scf.py:1-444).

rho = 2 sum_i |phi_i|^2;  F psi = T psi + V_ext psi + V_nl psi + V_H psi - K psi
with (K psi)(r) = sum_j phi_j(r) J[phi_j psi](r);
E = 2 sum_i <phi_i|h|phi_i> + E_H[rho] + E_x,  E_x = -sum_i <phi_i|K phi_i>.

Orbitals are exchanged with callers as (N, N, N, n) like the reference; they
are stored as contiguous (n, N, N, N) stacks (the (N, N, N, n) arrays handed
out are views).  ``precision`` selects the arithmetic of the fields:
"fp64" (everything double, the plain FP64 GPU path), "fp32" (fields, stencils
and FFTs in float; every reduction, Lanczos coefficient and the tridiagonal
eigensolve in double) and "mixed" (fp32 iterations, then an FP64 correction:
FP64 SCF iterations run until two consecutive ones differ by less than
tol).  By default the fp32 iterations of
"mixed" solve their eigenproblem inexactly: block Lanczos seeded with the
current orbitals, stopped once the Ritz residuals fall below a tolerance that
tightens with |dE| (inexact SCF), on the Fock operator with the compressed
exchange of the current orbitals; the FP64 correction keeps that scheme with
residuals below 0.05 sqrt(tol) (FP64_CORRECTION = "reference": the
reference's scalar Lanczos with its whole Krylov dimension and the exact
exchange).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import dataclasses
import functools
import math
import time
from dataclasses import dataclass, field
from typing import Callable

import torch

from . import _lib as L
from . import _native as nat
from .grid import Grid3D
from .hamiltonian import Hamiltonian
from .pseudopotentials import NonlocalPP
from .solvers.block_lanczos import MAX_BLOCK, block_lanczos
from .solvers.lanczos import LanczosGraph, default_krylov_dim, lanczos
from .solvers.poisson import get_plan

# SCF iterations replay their m native Lanczos steps from a HIP graph
# (recorded on the second iteration of a precision) unless switched off here
# or per call (use_graph=False).
GRAPH_LANCZOS = True

# Residual control of the float iterations of precision="mixed" with
# eigensolver="auto": iteration k stops its block Lanczos run once every Ritz
# residual ||F y - theta y|| is below
#     min(KRYLOV_RES_CAP, max(KRYLOV_RES_FLOOR, KRYLOV_RES_C * sqrt(|dE_{k-1}|)))
# (at most the whole Krylov dimension).  An eigenvector error of r/gap moves
# the density by as much and the energy by its square, so scaling with
# sqrt(|dE|) keeps the eigensolve error below what the SCF iteration itself
# changes.
#
# The same float iterations apply the exchange in compressed form
# (``compressed_exchange``): built once per iteration from the current
# orbitals (their n^2 exchange solves), applied in every Lanczos step as a
# rank-n_occ projection instead of n_occ Poisson solves per vector.  It is
# exact on the orbitals, so a self-consistent solution of the compressed
# problem is one of the exact problem; the FP64 iterations keep it too unless
# FP64_CORRECTION (below) asks for the reference's exact exchange and scalar
# Lanczos.
#
# Measured on H2O 128^3 from seeded random orbitals, tol 1e-7, errors against
# an FP64 solve to 1e-9 (the plain FP64 solve to 1e-7: 3.98 s; its own errors
# 1.7e-9 Ha total energy, 4.3e-5 Ha orbital energies, 1.7e-5 Ha/bohr forces):
#   C = 0.05, compressed exchange: 0.56 s, 3600 Krylov vectors, 1.2e-10 /
#     1.5e-5 / 9.2e-6;
#   C = 0.1: 0.43 s, 2708 vectors, 3.9e-9 / 4.4e-5 / 2.4e-5;
#   C = 0.2: 0.36 s, 2104 vectors, 6.4e-8 / 1.2e-4 / 6.6e-5;
#   C = 0.2 with the exact exchange: 0.68 s, 2180 vectors (the whole
#     dimension, 4800: 1.41 s);
# cap 1e-2 kept the first 13 iterations at the whole dimension; cap 3 ended
# the float phase early enough to need 6 FP64 iterations.
KRYLOV_RES_C = 0.05
KRYLOV_RES_FLOOR = 1e-4
KRYLOV_RES_CAP = 1.0
COMPRESSED_EXCHANGE = True
# The double iterations of precision="mixed" (eigensolver="auto"):
# "adaptive" keeps the float iterations' scheme (compressed exchange, block
# Lanczos seeded with the orbitals) with every Ritz residual below
# KRYLOV_RES_C * sqrt(tol), so the eigensolves add an energy error of order
# res^2 / gap, far below tol; "reference" runs the reference's scalar Lanczos
# over the whole Krylov dimension with the exact exchange.  On H2O 128^3 (tol
# 1e-7, the same float iterations) the double iteration took 16 ms
# (84 Krylov vectors) against 137 ms, with the same errors against an FP64
# solve to 1e-9 (energy 7e-10 / 1.0e-9 Ha, forces 1.3e-5 / 1.6e-5 Ha/bohr).
FP64_CORRECTION = "adaptive"


def to_stack(orbitals: torch.Tensor) -> torch.Tensor:
    """(N, N, N, n) -> contiguous (n, N, N, N) (no copy for views made by to_grid)."""
    if orbitals.dim() == 3:
        return nat.contig(orbitals.unsqueeze(0), "orbitals")
    return nat.contig(orbitals.permute(3, 0, 1, 2), "orbitals")


def to_grid(stack: torch.Tensor) -> torch.Tensor:
    """(n, N, N, N) -> (N, N, N, n) view."""
    return stack.permute(1, 2, 3, 0)


def _density(orbitals: torch.Tensor) -> torch.Tensor:
    """rho = 2 sum_i |phi_i|^2 (scf.py:61-63)."""
    return nat.density_mix(to_stack(orbitals), 1.0, None)


def exchange_apply(psi: torch.Tensor, orbitals: torch.Tensor, grid: Grid3D,
                   epsilon: float | None = None) -> torch.Tensor:
    """(K psi)(r) = sum_j phi_j(r) J[phi_j psi](r) (scf.py:66-85)."""
    orb = to_stack(orbitals).to(psi.dtype) if orbitals.dtype != psi.dtype else to_stack(orbitals)
    return get_plan(grid, epsilon, psi.dtype, psi.device).exchange_apply(psi, orb, 1.0)


def _nl_for_dtype(V_nl_apply, dtype):
    if isinstance(V_nl_apply, NonlocalPP):
        if V_nl_apply.dtype == dtype:
            return V_nl_apply
        from .pseudopotentials import nonlocal_plan
        from .geometry import Nuclei
        nuc = Nuclei(V_nl_apply.positions, torch.zeros(V_nl_apply.positions.shape[0], dtype=torch.float64))
        return nonlocal_plan(nuc, V_nl_apply.elements, V_nl_apply.grid, dtype)
    return None


@dataclass(frozen=True)
class RHFFock:
    """Closed-shell Fock operator (scf.py:114-158); duck-types as a Hamiltonian."""

    grid: Grid3D
    V_ext: torch.Tensor
    orbitals: torch.Tensor
    V_H: torch.Tensor | None = None
    boundary: str = "dirichlet"
    order: int = 4
    epsilon: float | None = None
    V_nl_apply: Callable | None = None
    # compressed exchange (xi, coef) from ``compressed_exchange``: the exchange
    # term is applied as -xi (dV xi^T psi), exact on the span of the orbitals
    ace: tuple | None = None

    def with_ace(self, xi: torch.Tensor, coef: torch.Tensor) -> "RHFFock":
        """The same operator with its exchange replaced by the compressed one
        (the effective potential and orbital caches carry over)."""
        f = dataclasses.replace(self, ace=(xi, coef))
        for k in ("_veff_cache", "_orb_cache"):
            if k in self.__dict__:
                f.__dict__[k] = self.__dict__[k]
        return f

    def density(self) -> torch.Tensor:
        return _density(self.orbitals)

    def hartree(self) -> torch.Tensor:
        if self.V_H is not None:
            return self.V_H
        from .solvers.poisson import hartree_potential
        return hartree_potential(self.density(), self.grid, self.epsilon)

    def _veff(self, dtype) -> torch.Tensor:
        cache = self.__dict__.setdefault("_veff_cache", {})
        v = cache.get(dtype)
        if v is None:
            v = nat.convert(self.V_ext, dtype)
            VH = self.hartree()
            nat.axpby(1.0, VH if VH.dtype == dtype else nat.convert(VH, dtype), 1.0, v)
            cache[dtype] = v
        return v

    def _orb_stack(self, dtype) -> torch.Tensor:
        cache = self.__dict__.setdefault("_orb_cache", {})
        o = cache.get(dtype)
        if o is None:
            o = to_stack(self.orbitals)
            if o.dtype != dtype:
                o = nat.convert(o, dtype)
            cache[dtype] = o
        return o

    @contextlib.contextmanager
    def _handle(self, dtype):
        V = self._veff(dtype)
        orb = self._orb_stack(dtype)
        plan = get_plan(self.grid, self.epsilon, dtype, V.device)
        nl = _nl_for_dtype(self.V_nl_apply, dtype)
        h = C.c_void_p()
        n3 = self.grid.N ** 3
        L.check(L.lib().gato_fock_create(
            L.dtype_code(dtype), self.grid.N, self.grid.h, nat.check_order(self.order),
            nat.boundary_code(self.boundary), V.data_ptr(), plan.handle, orb.data_ptr(), n3,
            orb.shape[0], nl.handle if nl is not None else None, C.byref(h)))
        try:
            if self.ace is not None:
                xi, coef = self.ace
                if xi.dtype != dtype:
                    raise ValueError(f"compressed exchange is {xi.dtype}, operator applied in "
                                     f"{dtype}")
                L.check(L.lib().gato_fock_set_ace(h, xi.data_ptr(), xi.shape[0],
                                                  coef.data_ptr()))
            yield h
        finally:
            L.lib().gato_fock_destroy(h)

    @property
    def _native_fock(self):
        if self.V_nl_apply is not None and not isinstance(self.V_nl_apply, NonlocalPP):
            return None
        return self._handle

    def apply(self, psi: torch.Tensor) -> torch.Tensor:
        """F psi (scf.py:146-158)."""
        psi = nat.contig(psi, "psi")
        out = torch.empty_like(psi)
        native_nl = self.V_nl_apply is None or isinstance(self.V_nl_apply, NonlocalPP)
        if native_nl:
            with self._handle(psi.dtype) as h:
                L.check(L.lib().gato_fock_apply(h, psi.data_ptr(), out.data_ptr(), None, None,
                                                L.workspace(psi.device).data_ptr(),
                                                L.stream_ptr()))
            return out
        # arbitrary callable for V_nl: native Fock without it, plus the callable
        with contextlib.ExitStack():
            V = self._veff(psi.dtype)
            orb = self._orb_stack(psi.dtype)
            plan = get_plan(self.grid, self.epsilon, psi.dtype, psi.device)
            plan.exchange_apply(psi, orb, -1.0, out=out)
            nat.stencil(psi, self.grid.h, self.order, self.boundary, -0.5, V=V, out=out,
                        accumulate=True)
            extra = self.V_nl_apply(psi)
            nat.axpby(1.0, extra.to(psi.dtype).contiguous(), 1.0, out)
        return out


# ---------------------------------------------------------------------------
# Energy decomposition (scf.py:165-286)
# ---------------------------------------------------------------------------

def ace_exchange_energy(coef: torch.Tensor, n: int) -> float:
    """E_x = -sum_i <phi_i|K phi_i> of the orbitals a compressed exchange was
    built on: coef[:n*n] holds M = dV phi^T (K phi) right after the build
    (before an operator apply reuses the buffer), and <phi_i|K phi_i> =
    sum_j (ij|ij) (scf.py:225-229), so -trace(M) is the pair-integral sum of
    the energy.  Summed on the host in orbital order."""
    m = coef[: n * n].cpu().numpy()
    ex = 0.0
    for i in range(n):
        ex -= float(m[i * n + i])
    return ex


def compressed_exchange(fock: RHFFock, dtype: torch.dtype) -> tuple:
    """Adaptively compressed exchange (L. Lin, J. Chem. Theory Comput. 12,
    2242 (2016)) of ``fock``'s exact exchange in ``dtype``: W = K phi over
    its orbitals, M = <phi, W> = R^T R, xi = W R^{-1}, so that
    xi <xi, psi> = W M^{-1} <W, psi> equals K psi on the span of the
    orbitals.  Returns (xi, coef) for ``RHFFock.with_ace``.  No reference
    counterpart: the reference applies the exact exchange in every Lanczos
    step (scf.py:146-158)."""
    if fock.V_nl_apply is not None and not isinstance(fock.V_nl_apply, NonlocalPP):
        raise ValueError("compressed exchange needs the native Fock operator")
    orb = fock._orb_stack(dtype)
    n = orb.shape[0]
    xi = torch.empty_like(orb)
    coef = torch.empty(max(4 * n * n, 8 * n), dtype=torch.float64, device=orb.device)
    with fock._handle(dtype) as h:
        L.check(L.lib().gato_ace_build(h, xi.data_ptr(), coef.data_ptr(),
                                       L.workspace(orb.device).data_ptr(), L.stream_ptr()))
    return xi, coef


@dataclass(frozen=True)
class RHFEnergyTerms:
    """The energy decomposition in float64 on the host: the per-orbital and
    per-pair integrals come from the HIP kernels, their weighted sums are
    formed here (a few doubles, one copy each)."""
    kinetic: float    # 2 sum_i <phi_i|T|phi_i>
    v_ext: float      # 2 sum_i <phi_i|V_ext|phi_i>
    v_nl: float       # 2 sum_i <phi_i|V_nl|phi_i>
    hartree: float    # E_H[rho]
    exchange: float   # E_x

    @property
    def one_body(self):
        return self.kinetic + self.v_ext + self.v_nl

    @property
    def total(self):
        return self.kinetic + self.v_ext + self.v_nl + self.hartree + self.exchange


def _energy_terms(orb: torch.Tensor, V_ext: torch.Tensor, grid: Grid3D, boundary: str,
                  order: int, epsilon, V_nl_apply, ace_m: torch.Tensor | None = None
                  ) -> RHFEnergyTerms:
    """orb: (n, N, N, N) stack in the working dtype.  ``ace_m``: the n x n
    device matrix dV phi^T (K phi) of a compressed exchange built on these
    orbitals (row-major), whose trace gives E_x; without it the pair integrals
    do.  Every integral stays on the device until one copy brings them all to
    the host, where their weighted sums are formed in a fixed order."""
    dt = orb.dtype
    n = orb.shape[0]
    n3 = grid.N ** 3
    V = V_ext if V_ext.dtype == dt else nat.convert(V_ext, dt)
    Tphi = nat.stencil(orb, grid.h, order, boundary, -0.5)
    parts = [nat.dot_batched(orb, Tphi, n, n3, n3, scale=grid.dV),
             nat.dot_batched(orb, orb, n, n3, n3, C_=nat.contig(V), scale=grid.dV)]
    Vnl = 0.0
    nl = None
    if V_nl_apply is not None:
        nl = _nl_for_dtype(V_nl_apply, dt)
        if nl is not None:
            parts.append(nl.expectation(orb))
        else:
            Vnl = sum(float(nat.dot_batched(orb[i], V_nl_apply(orb[i]).to(dt).contiguous(), 1, 0,
                                            0, scale=grid.dV)[0]) for i in range(n))
    plan = get_plan(grid, epsilon, dt, orb.device)
    rho = nat.density_mix(orb, 1.0, None)
    parts.append(plan.hartree_energy(rho)[:1])
    parts.append(ace_m.reshape(-1)[: n * n] if ace_m is not None else plan.pair_integrals(orb))
    host = torch.cat([p.reshape(-1).to(torch.float64) for p in parts]).cpu().numpy()
    T, Ve, k = host[:n], host[n:2 * n], 2 * n
    if nl is not None:
        Vnl = float(host[k:k + n].sum())
        k += n
    EH = float(host[k])
    rest = host[k + 1:]
    Ex = 0.0
    if ace_m is not None:
        # E_x = -sum_i <phi_i|K phi_i> (scf.py:225-229), the trace of dV phi^T K phi
        for i in range(n):
            Ex -= float(rest[i * n + i])
    else:
        # E_x = -sum_ij (ij|ij): the pair integrals for i <= j, diagonal pairs
        # once and the others twice
        idx = 0
        for i in range(n):
            for j in range(i, n):
                Ex -= (1.0 if i == j else 2.0) * float(rest[idx])
                idx += 1
    return RHFEnergyTerms(2.0 * float(T.sum()), 2.0 * float(Ve.sum()), 2.0 * Vnl, EH, Ex)


def decompose_rhf_energy(orbitals: torch.Tensor, V_ext: torch.Tensor, grid: Grid3D, *,
                         boundary: str = "dirichlet", order: int = 4, epsilon: float | None = None,
                         V_nl_apply: Callable | None = None) -> RHFEnergyTerms:
    """Every term of the closed-shell RHF energy (scf.py:234-265).  E_x uses the
    n(n+1)/2 distinct pair densities: -(sum_i (ii|ii) + 2 sum_{i<j} (ij|ij))."""
    nat.check_order(order)
    nat.boundary_code(boundary)
    return _energy_terms(to_stack(orbitals), V_ext, grid, boundary, order, epsilon, V_nl_apply)


def rhf_energy(orbitals, V_ext, grid, boundary="dirichlet", order=4, epsilon=None,
               V_nl_apply=None) -> torch.Tensor:
    """decompose_rhf_energy(...).total (scf.py:268-286)."""
    return decompose_rhf_energy(orbitals, V_ext, grid, boundary=boundary, order=order,
                                epsilon=epsilon, V_nl_apply=V_nl_apply).total


def default_initial_vector(grid: Grid3D, dtype=torch.float64) -> torch.Tensor:
    """(e^{-4r} + e^{-r}) (1 + 0.1(x+y+z) + 0.02(xy+yz+xz)) (scf.py:310-320)."""
    X, Y, Z = grid.coords()
    r = torch.sqrt(X * X + Y * Y + Z * Z + 1e-6)
    pert = 0.1 * (X + Y + Z) + 0.02 * (X * Y + Y * Z + X * Z)
    return ((torch.exp(-4.0 * r) + torch.exp(-r)) * (1.0 + pert)).to(dtype).contiguous()


def default_initial_block(grid: Grid3D, b: int, dtype=torch.float64) -> torch.Tensor:
    """b start vectors for block Lanczos, (N, N, N, b): the radial part of
    ``default_initial_vector`` times 1, x, y, z, xy, yz, zx, x^2 - y^2, so
    every symmetry species of a molecule enters with its own direction (a
    degenerate pair needs two)."""
    X, Y, Z = grid.coords()
    r = torch.sqrt(X * X + Y * Y + Z * Z + 1e-6)
    base = torch.exp(-4.0 * r) + torch.exp(-r)
    f = [torch.ones_like(X), X, Y, Z, X * Y, Y * Z, Z * X, X * X - Y * Y]
    return torch.stack([(base * f[c]).to(dtype) for c in range(b)], dim=-1)


def _default_initial_orbitals(V_ext, grid, n_occ, boundary, order, lanczos_iters,
                              dtype=torch.float64, eigensolver: str = "lanczos") -> torch.Tensor:
    """Core-Hamiltonian guess (scf.py:289-322); returns an (n, N, N, N) stack.
    With block Lanczos the guess is itself a block run from
    ``default_initial_block``: a scalar run would leave every degenerate
    level with a single direction for all later iterations to inherit."""
    H = Hamiltonian(grid=grid, V=V_ext, boundary=boundary, order=order)
    if eigensolver == "block_lanczos":
        res = block_lanczos(H, default_initial_block(grid, n_occ, dtype), n_iters=lanczos_iters,
                            n_eigenstates=n_occ)
        return res.states
    res = lanczos(H, default_initial_vector(grid, dtype), n_iters=lanczos_iters,
                  n_eigenstates=n_occ)
    return res.states


@dataclass
class SCFResult:
    orbitals: torch.Tensor            # (N, N, N, n_occ)
    energy: float
    orbital_energies: torch.Tensor    # (n_occ,)
    n_iters: int
    converged: bool
    energy_history: list
    precision_history: list = field(default_factory=list)
    iteration_seconds: list = field(default_factory=list)
    krylov_history: list = field(default_factory=list)  # Krylov vectors per iteration


def _sum_stack(orb: torch.Tensor) -> torch.Tensor:
    """sum_i phi_i as one field (scf.py:417), via the Ritz-combine kernel."""
    n = orb.shape[0]
    n3 = orb[0].numel()
    S = torch.tensor([1.0] * n, dtype=torch.float64, device=orb.device)
    out = torch.empty(orb.shape[1:], dtype=orb.dtype, device=orb.device)
    L.check(L.lib().gato_ritz_combine(L.dtype_code(orb.dtype), n3, n, 1, orb.data_ptr(), n3,
                                      S.data_ptr(), out.data_ptr(), n3, L.stream_ptr()))
    return out


_PRECISIONS = {"fp64": torch.float64, "fp32": torch.float32, "mixed": torch.float32}
_EIGENSOLVERS = ("lanczos", "block_lanczos")


def krylov_res_tol(dE: float) -> float:
    """Ritz-residual tolerance of an adaptive float iteration after an energy change dE."""
    if not math.isfinite(dE):
        return KRYLOV_RES_CAP
    return min(KRYLOV_RES_CAP, max(KRYLOV_RES_FLOOR, KRYLOV_RES_C * math.sqrt(abs(dE))))


class SCFIterator:
    """One closed-shell SCF iteration per ``step()`` (the body of scf.py:394-435):
    Hartree of the mixed density, Fock with the current orbitals for exchange,
    Lanczos seeded by sum_i phi_i, linear density mixing, energy of the new
    orbitals.  ``step()`` returns the energy (one host sync, like the
    reference's ``float(rhf_energy(...))``).

    eigensolver="block_lanczos" replaces the scalar Lanczos run by block
    Lanczos seeded with the n_occ current orbitals themselves (n_occ <= 8),
    with about ``lanczos_iters`` Krylov vectors: the remedy for degenerate
    occupied levels that lanczos.py:83-107 names.

    use_graph (scalar Lanczos with the native Fock operator): the effective
    potential, the exchange orbitals and the seed are written in place into
    buffers the iterator keeps, and the m Lanczos steps run from a HIP graph
    recorded on the second iteration of each precision (``LanczosGraph``);
    results equal the eager path bit for bit.

    ``res_tol`` (block Lanczos only; set between steps): stop each block
    Lanczos run once its Ritz residuals are below it (``block_lanczos``'s
    adaptive Krylov dimension); ``krylov_used`` is the last run's vector
    count.  ``ace`` (set between steps; not with the graph): the eigensolver's
    Fock operator takes the compressed exchange of the current orbitals
    (``compressed_exchange``) instead of the exact exchange."""

    def __init__(self, V_ext: torch.Tensor, grid: Grid3D, n_occ: int, *,
                 initial_orbitals: torch.Tensor | None = None, mixing: float = 1.0,
                 boundary: str = "dirichlet", order: int = 4, lanczos_iters: int | None = None,
                 epsilon: float | None = None, V_nl_apply: Callable | None = None,
                 dtype: torch.dtype = torch.float64, eigensolver: str = "lanczos",
                 use_graph: bool | None = None):
        nat.check_order(order)
        nat.boundary_code(boundary)
        if eigensolver not in _EIGENSOLVERS:
            raise ValueError(f"eigensolver must be one of {_EIGENSOLVERS}")
        if eigensolver == "block_lanczos" and not 1 <= n_occ <= MAX_BLOCK:
            raise ValueError(f"block_lanczos seeds one block with the {n_occ} occupied orbitals; "
                             f"at most {MAX_BLOCK} are supported")
        self.eigensolver = eigensolver
        self.res_tol: float | None = None
        self.ace = False
        self.krylov_used = 0
        self.grid, self.n_occ, self.mixing = grid, int(n_occ), float(mixing)
        self.boundary, self.order, self.epsilon, self.V_nl_apply = boundary, order, epsilon, V_nl_apply
        self.m = default_krylov_dim(grid.N, n_occ) if lanczos_iters is None else int(lanczos_iters)
        V64 = V_ext if V_ext.dtype == torch.float64 else nat.convert(V_ext, torch.float64)
        self._V = {torch.float64: nat.contig(V64)}
        self.dtype = dtype
        if initial_orbitals is None:
            self.orb = _default_initial_orbitals(self.V_as(dtype), grid, self.n_occ, boundary,
                                                 order, self.m, dtype, eigensolver)
        else:
            orb = to_stack(initial_orbitals)
            self.orb = orb if orb.dtype == dtype else nat.convert(orb, dtype)
        self.rho = nat.density_mix(self.orb, 1.0, None)
        self.orbital_energies = torch.zeros(self.n_occ, dtype=torch.float64)
        self.lanczos_converged = False
        self._native = V_nl_apply is None or isinstance(V_nl_apply, NonlocalPP)
        self._graph_ok = GRAPH_LANCZOS if use_graph is None else bool(use_graph)
        self.use_graph = self._graph_ok and self._native and eigensolver == "lanczos"
        self._graphs: dict = {}
        self._nl: dict = {}
        # (orbitals, operator with its compressed exchange) prepared by the
        # previous step for these orbitals
        self._next: tuple | None = None

    def set_eigensolver(self, eigensolver: str) -> None:
        """Switch the eigensolver of later steps (graph replay follows use_graph)."""
        if eigensolver not in _EIGENSOLVERS:
            raise ValueError(f"eigensolver must be one of {_EIGENSOLVERS}")
        if eigensolver == "block_lanczos" and not 1 <= self.n_occ <= MAX_BLOCK:
            raise ValueError(f"block_lanczos needs 1..{MAX_BLOCK} occupied orbitals")
        self.eigensolver = eigensolver
        self.use_graph = self._graph_ok and self._native and eigensolver == "lanczos"

    def V_as(self, dtype: torch.dtype) -> torch.Tensor:
        if dtype not in self._V:
            self._V[dtype] = nat.convert(self._V[torch.float64], dtype)
        return self._V[dtype]

    def energy_terms(self, ace_m: torch.Tensor | None = None) -> RHFEnergyTerms:
        return _energy_terms(self.orb, self.V_as(self.dtype), self.grid, self.boundary, self.order,
                             self.epsilon, self.V_nl_apply, ace_m)

    def promote(self, dtype: torch.dtype) -> None:
        """Switch the working precision (orbitals and mixed density are converted)."""
        if dtype != self.dtype:
            self.orb = nat.convert(self.orb, dtype)
            self.rho = nat.convert(self.rho, dtype)
            self.dtype = dtype
            self._next = None
            for g in self._graphs.values():
                g.close()
            self._graphs.clear()

    def _graph_lanczos(self, dt: torch.dtype, V_H: torch.Tensor):
        """The Lanczos run of step() from the recorded graph of this precision."""
        g = self._graphs.get(dt)
        if g is None:
            nl = self._nl.get(dt)
            if nl is None and self.V_nl_apply is not None:
                nl = self._nl[dt] = _nl_for_dtype(self.V_nl_apply, dt)
            g = LanczosGraph(self.grid, dt, self.n_occ, self.m, self.n_occ, self.orb.device,
                             self.order, self.boundary, self.epsilon, nl)
            self._graphs[dt] = g
        # the same effective potential as RHFFock: V_ext in dt, plus V_H
        g.veff.copy_(self.V_as(dt))
        nat.axpby(1.0, V_H if V_H.dtype == dt else nat.convert(V_H, dt), 1.0, g.veff)
        g.orb.copy_(self.orb)
        g.seed.copy_(_sum_stack(self.orb))                                          # scf.py:417
        return g.run()

    def _fock(self, dt: torch.dtype, orb: torch.Tensor, rho: torch.Tensor) -> RHFFock:
        plan = get_plan(self.grid, self.epsilon, dt, orb.device)
        V_H = plan.hartree_potential(rho)                                            # scf.py:399
        return RHFFock(grid=self.grid, V_ext=self.V_as(dt), orbitals=to_grid(orb), V_H=V_H,
                       boundary=self.boundary, order=self.order, epsilon=self.epsilon,
                       V_nl_apply=self.V_nl_apply)

    def step(self) -> float:
        dt = self.dtype
        nxt, self._next = self._next, None
        with_ace = self.ace and not self.use_graph
        if with_ace and nxt is not None and nxt[0] is self.orb and nxt[1].ace[0].dtype == dt:
            fock = nxt[1]  # built by the previous step from these orbitals and this density
        else:
            fock = self._fock(dt, self.orb, self.rho)
            if with_ace:
                fock = fock.with_ace(*compressed_exchange(fock, dt))
        V_H = fock.V_H
        if self.eigensolver == "block_lanczos":
            res = block_lanczos(fock, to_grid(self.orb), n_iters=self.m,
                                n_eigenstates=self.n_occ, res_tol=self.res_tol)
        elif self.use_graph:
            res = self._graph_lanczos(dt, V_H)                                       # scf.py:418
        else:
            seed = _sum_stack(self.orb)                                              # scf.py:417
            res = lanczos(fock, seed, n_iters=self.m, n_eigenstates=self.n_occ)      # scf.py:418
        self.lanczos_converged = res.converged
        self.krylov_used = res.n_iters
        new = res.states
        self.orbital_energies = res.eigenvalues
        self.rho = nat.density_mix(new, self.mixing, self.rho)                       # scf.py:422-423
        self.orb = new
        ace_m = None
        if with_ace:
            # the next iteration's operator now: its compressed exchange is built
            # on the new orbitals, so dV phi^T K phi gives the energy's exchange
            # term (scf.py:225-229) without the n(n+1)/2 pair solves (read
            # before the next apply reuses coef)
            nf = self._fock(dt, new, self.rho)
            xi, coef = compressed_exchange(nf, dt)
            ace_m = coef[: self.n_occ ** 2]
            self._next = (new, nf.with_ace(xi, coef))
        return float(self.energy_terms(ace_m).total)                                 # scf.py:426-429


def scf_rhf(V_ext: torch.Tensor, grid: Grid3D, n_occ: int, *,
            initial_orbitals: torch.Tensor | None = None, max_iters: int = 50, tol: float = 1e-6,
            mixing: float = 1.0, boundary: str = "dirichlet", order: int = 4,
            lanczos_iters: int | None = None, epsilon: float | None = None,
            V_nl_apply: Callable | None = None, precision: str = "fp64",
            switch_tol: float | None = None, eigensolver: str = "auto",
            on_iteration: Callable | None = None, use_graph: bool | None = None) -> SCFResult:
    """SCF loop with linear density mixing (scf.py:335-444).

    precision="mixed": float iterations until |dE| < switch_tol (default
    max(tol, 3e-7)), or until the float energies stop improving once |dE| is
    within 100x of switch_tol (three iterations in a row without a new
    smallest |dE|: the float rounding floor of a larger grid can sit above
    switch_tol).  Then the iterations continue in double until two
    consecutive double iterations differ by less than tol, the stopping rule
    of the plain FP64 loop applied to double iterates only (so at least two
    of them run).  The float energies carry the float evaluation's own
    offset (2.6e-5 Ha on H2O 128^3), so a double energy is never compared
    with a float one; stopping after one double iteration measured against
    the double re-evaluation of the float orbitals left the energy 4e-8 Ha
    from an FP64 solve to 1e-9 on H2O 128^3 (tol 1e-7).
    eigensolver="block_lanczos", use_graph: see ``SCFIterator``.

    eigensolver="auto" (the default) is the reference's scalar Lanczos with
    its whole Krylov dimension, except in the float iterations of
    precision="mixed" with at most 8 occupied orbitals: those run block
    Lanczos seeded with the current orbitals and stop it once the Ritz
    residuals are below ``krylov_res_tol(|dE|)`` of the previous iteration
    (inexact eigensolves while the SCF is far from its fixed point, tighter
    ones near it).  The double iterations keep that scheme with residuals
    below ``KRYLOV_RES_C * sqrt(tol)`` (``FP64_CORRECTION = "adaptive"``) or
    run the scalar Lanczos of the reference with the exact exchange
    (``"reference"``); the fixed point is the same: the compressed exchange
    equals the exact one on the orbitals.
    "lanczos" or "block_lanczos" run that solver with the whole Krylov
    dimension in every iteration.
    """
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
    if eigensolver not in _EIGENSOLVERS + ("auto",):
        raise ValueError(f"eigensolver must be one of {_EIGENSOLVERS + ('auto',)}")
    if switch_tol is None:
        switch_tol = max(tol, 3e-7)
    adaptive = eigensolver == "auto" and precision == "mixed" and 1 <= n_occ <= MAX_BLOCK
    first = "block_lanczos" if adaptive else ("lanczos" if eigensolver == "auto" else eigensolver)
    it_ = SCFIterator(V_ext, grid, n_occ, initial_orbitals=initial_orbitals, mixing=mixing,
                      boundary=boundary, order=order, lanczos_iters=lanczos_iters,
                      epsilon=epsilon, V_nl_apply=V_nl_apply, dtype=_PRECISIONS[precision],
                      eigensolver=first, use_graph=use_graph)
    native = V_nl_apply is None or isinstance(V_nl_apply, NonlocalPP)
    it_.ace = adaptive and native and COMPRESSED_EXCHANGE
    history, prec_hist, secs, kry = [], [], [], []
    E_prev = math.inf
    E = math.inf
    dE_last = math.inf   # the previous iteration's |dE| (adaptive Krylov dimension)
    converged = False
    it = 0
    best32, stall32 = math.inf, 0
    for it in range(1, max_iters + 1):
        if adaptive and it_.dtype == torch.float32:
            it_.res_tol = krylov_res_tol(dE_last)
        t0 = time.perf_counter()
        E = it_.step()
        secs.append(time.perf_counter() - t0)
        kry.append(it_.krylov_used)
        history.append(E)
        prec_hist.append("fp32" if it_.dtype == torch.float32 else "fp64")
        if on_iteration is not None:
            on_iteration(it, E, prec_hist[-1], secs[-1])
        dE = abs(E - E_prev)
        dE_last = dE
        if precision == "mixed" and it_.dtype == torch.float32:
            if dE < best32:
                best32, stall32 = dE, 0
            elif best32 < 100.0 * switch_tol:
                stall32 += 1
            if dE < switch_tol or stall32 >= 3:
                # FP64 correction: promote; convergence is judged between
                # consecutive double iterations, as in the plain FP64 loop, so
                # the first one is not compared with a float energy
                it_.promote(torch.float64)
                if adaptive and FP64_CORRECTION == "adaptive":
                    it_.res_tol = KRYLOV_RES_C * math.sqrt(tol)
                elif adaptive:  # the reference's solver and exchange for the double iterations
                    it_.res_tol = None
                    it_.ace = False
                    it_.set_eigensolver("lanczos")
                E_prev = math.inf
            else:
                E_prev = E
            continue
        if dE < tol:                                                                 # scf.py:432-435
            converged = True
            break
        E_prev = E
    return SCFResult(to_grid(it_.orb), E, it_.orbital_energies, it, converged, history,
                     prec_hist, secs, kry)
