"""HGH norm-conserving pseudopotentials (reference: pseudopotentials.py:1-405).

Local part V_loc(r) = -Z_ion erf(r / (sqrt2 r_loc)) / r
                      + exp(-r^2 / (2 r_loc^2)) sum_i C_i (r/r_loc)^{2(i-1)},
sampled by a HIP kernel in double.  Nonlocal part
V_nl psi = sum_k sum_lm sum_ij p_i h_ij <p_j|psi>, with the projectors sampled
once per geometry on compact per-atom boxes (``NonlocalPP``) and applied by
two small kernels per atom.

The parameter tables below are the published HGH-1998 (LDA) values, the same
numbers the reference tabulates at pseudopotentials.py:122-185 (a physical
constant table has to be identical to give identical results).
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass

import torch

from . import _lib as L
from . import _native as nat
from .geometry import Nuclei
from .grid import Grid3D


@dataclass(frozen=True)
class HGHProjector:
    """One l-channel: width r_l and symmetric coupling h_ij (Ha)."""

    ell: int
    r: float
    h: tuple

    @property
    def n(self) -> int:
        return len(self.h)


@dataclass(frozen=True)
class HGHParams:
    symbol: str
    Z_ion: float
    r_loc: float
    c: tuple
    projectors: tuple = ()


HYDROGEN_HGH = HGHParams("H", 1.0, 0.2, (-4.0663326, 0.6778322, 0.0, 0.0), ())
OXYGEN_HGH = HGHParams("O", 6.0, 0.24762086, (-16.5803180, 2.39570952, 0.0, 0.0),
                       (HGHProjector(0, 0.22178614, ((18.26691718,),)),))
LITHIUM_HGH = HGHParams("Li", 1.0, 0.78755305, (-1.89261247, 0.28605968, 0.0, 0.0),
                        (HGHProjector(0, 0.66637518, ((1.85881111,),)),
                         HGHProjector(1, 1.07930561, ((-0.00589504,),))))
CHLORINE_HGH = HGHParams("Cl", 7.0, 0.41, (-6.86475431, 0.0, 0.0, 0.0),
                         (HGHProjector(0, 0.33820832, ((9.06223968, -1.96193036),
                                                       (-1.96193036, 5.06568240))),
                          HGHProjector(1, 0.37613709, ((4.46587640,),))))

HGH_TABLE = {"H": HYDROGEN_HGH, "Li": LITHIUM_HGH, "O": OXYGEN_HGH, "Cl": CHLORINE_HGH}


def lookup(symbol: str) -> HGHParams:
    """pseudopotentials.py:188-194 (KeyError for unknown elements)."""
    if symbol not in HGH_TABLE:
        raise KeyError(f"No HGH parameters for {symbol!r}. Add them to gato_amd.pseudopotentials "
                       f"from Hartwigsen-Goedecker-Hutter 1998 table I.")
    return HGH_TABLE[symbol]


def total_valence_charge(elements) -> float:
    return float(sum(p.Z_ion for p in elements))


def local_centers(nuclei: Nuclei, elements) -> list:
    if len(elements) != nuclei.n:
        raise ValueError(f"HGH parameter list length {len(elements)} != number of nuclei {nuclei.n}")
    P = nuclei.positions.detach().cpu().tolist()
    out = []
    for k, p in enumerate(elements):
        out.append([L.CENTER_HGH_LOCAL, P[k][0], P[k][1], P[k][2], p.Z_ion, p.r_loc, *p.c])
    return out


def hgh_local_potential(nuclei: Nuclei, elements, grid: Grid3D, dtype=torch.float64) -> torch.Tensor:
    """Sum of HGH local parts over nuclei (pseudopotentials.py:230-253)."""
    return nat.local_potential(grid.N, grid.L, local_centers(nuclei, elements), dtype)


def projector_channels(elements):
    """(atom, l, m, i, r_l) per channel, in the reference's loop order
    (pseudopotentials.py:338-356), plus the block-diagonal coupling matrix."""
    chans = []
    blocks = []
    for k, p in enumerate(elements):
        for proj in p.projectors:
            for m in range(-proj.ell, proj.ell + 1):
                first = len(chans)
                for i in range(proj.n):
                    chans.append((k, proj.ell, m, i + 1, proj.r))
                blocks.append((first, proj.h))
    n = len(chans)
    H = [[0.0] * n for _ in range(n)]
    for first, h in blocks:
        for a in range(len(h)):
            for b in range(len(h)):
                H[first + a][first + b] = float(h[a][b])
    return chans, H


class NonlocalPP:
    """Device plan of the separable nonlocal part at one geometry.

    Callable as ``V_nl(psi)`` (the ``V_nl_apply`` protocol of scf.py:98).
    ``cutoff_tol`` bounds the relative projector value dropped outside each
    atom's box (the reference samples the whole grid).
    """

    def __init__(self, nuclei: Nuclei, elements, grid: Grid3D, dtype=torch.float64,
                 cutoff_tol: float | None = None):
        if len(elements) != nuclei.n:
            raise ValueError(f"HGH parameter list length {len(elements)} != number of nuclei {nuclei.n}")
        self.grid, self.dtype, self.elements = grid, dtype, tuple(elements)
        self.positions = nuclei.positions.detach().cpu().to(torch.float64).clone()
        if cutoff_tol is None:
            cutoff_tol = 1e-14 if dtype == torch.float64 else 1e-9
        chans, H = projector_channels(elements)
        self.n_proj = len(chans)
        self._h = C.c_void_p()
        pos = L.double_array(self.positions.reshape(-1).tolist())
        L.check(L.lib().gato_nl_plan_create(
            L.dtype_code(dtype), grid.N, grid.L, nuclei.n, pos, self.n_proj,
            L.int_array([c[0] for c in chans]), L.int_array([c[1] for c in chans]),
            L.int_array([c[2] for c in chans]), L.int_array([c[3] for c in chans]),
            L.double_array([c[4] for c in chans]), L.double_array([v for row in H for v in row]),
            float(cutoff_tol), C.byref(self._h), L.stream_ptr()))

    @property
    def handle(self):
        return self._h

    @property
    def n_points(self) -> int:
        return int(L.lib().gato_nl_plan_points(self._h))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            L.lib().gato_nl_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, psi):
        psi = nat.contig(psi, "psi")
        if psi.dtype != self.dtype:
            raise ValueError(f"psi dtype {psi.dtype} does not match the plan ({self.dtype})")
        return psi

    def __call__(self, psi: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """V_nl psi for a field or a (k, N, N, N) stack; adds into ``out`` if given."""
        psi = self._check(psi)
        n3 = self.grid.N ** 3
        if out is None:
            out = torch.zeros_like(psi)
        L.check(L.lib().gato_nl_apply(self._h, psi.numel() // n3, psi.data_ptr(), n3,
                                      out.data_ptr(), n3, L.workspace(psi.device).data_ptr(),
                                      L.stream_ptr()))
        return out

    def expectation(self, psi: torch.Tensor) -> torch.Tensor:
        """<psi_v|V_nl|psi_v> for each field of a stack -> (k,) float64."""
        psi = self._check(psi)
        n3 = self.grid.N ** 3
        k = psi.numel() // n3
        E = torch.empty(k, dtype=torch.float64, device=psi.device)
        L.check(L.lib().gato_nl_expectation(self._h, k, psi.data_ptr(), n3, E.data_ptr(),
                                            L.workspace(psi.device).data_ptr(), L.stream_ptr()))
        return E

    def grad(self, psi: torch.Tensor, occ: float = 2.0) -> torch.Tensor:
        """d/dR sum_v occ <psi_v|V_nl(R)|psi_v> -> (K, 3) float64 device tensor."""
        psi = self._check(psi)
        n3 = self.grid.N ** 3
        K = self.positions.shape[0]
        g = torch.empty(3 * K, dtype=torch.float64, device=psi.device)
        L.check(L.lib().gato_nl_grad(self._h, psi.numel() // n3, psi.data_ptr(), n3, float(occ),
                                     g.data_ptr(), L.workspace(psi.device).data_ptr(),
                                     L.stream_ptr()))
        return g.view(K, 3)


_NL_CACHE: dict = {}
_NL_LOCK = threading.Lock()  # host threads of a batch share the cache


def nonlocal_plan(nuclei: Nuclei, elements, grid: Grid3D, dtype=torch.float64) -> NonlocalPP:
    """The cached plan for this geometry on the current stream (its tables are
    written on the stream that builds it).  A plan dropped from the cache is
    freed when its last user releases it."""
    key = (tuple(nuclei.positions.detach().cpu().reshape(-1).tolist()), tuple(elements), grid, dtype,
           L.stream_ptr())
    with _NL_LOCK:
        p = _NL_CACHE.get(key)
    if p is None:
        p = NonlocalPP(nuclei, elements, grid, dtype)
        with _NL_LOCK:
            while len(_NL_CACHE) >= 16:
                L.retire(_NL_CACHE.pop(next(iter(_NL_CACHE))))
            _NL_CACHE[key] = p
    return p


def hgh_nonlocal_apply(psi: torch.Tensor, nuclei: Nuclei, elements, grid: Grid3D) -> torch.Tensor:
    """V_nl psi (pseudopotentials.py:360-383)."""
    return nonlocal_plan(nuclei, elements, grid, psi.dtype)(psi)


def hgh_apply(psi, nuclei, elements, grid, V_local=None):
    """V_loc psi + V_nl psi (pseudopotentials.py:390-405)."""
    if V_local is None:
        V_local = hgh_local_potential(nuclei, elements, grid, psi.dtype)
    out = V_local * psi
    return nonlocal_plan(nuclei, elements, grid, psi.dtype)(psi, out=out)
