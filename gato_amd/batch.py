"""Independent molecules on uneven grids, solved together on one GPU.

The workload is a seeded draw of small closed-shell molecules. Each molecule is:
- all-electron H2 (h2.py:67-102), or LiH or HCl with HGH pseudopotentials
  (lih.py:126-141, hcl.py:157-173);
- on its own grid, N in {64, 80, 96, 128};
- at a bond length uniform within +-15 % of equilibrium (R_e = 1.40, 3.02 and
  2.41 bohr: h2.py:22, lih.py:22, hcl.py:28);
- in the reference drivers' default box and mixing (h2.py:211-218,
  lih.py:260-267, hcl.py:294-301).

HCl at N >= 80 takes 200 Krylov vectors, the count hcl.py:40-51 names for its
nearly degenerate 3p manifold.

Execution: S slots share a queue of molecules, largest first. Each slot is a
host thread with its own HIP stream at the default priority, its own
reduction workspace and its own Poisson and nonlocal plans (all keyed by the
current stream). A slot runs one molecule's whole SCF on its stream, then
takes the next. Every kernel of a molecule is stream-ordered and the
reductions sum in a fixed order, so each molecule's result does not depend on
S or on what shares the GPU with it. Small grids leave most of the chip idle
between launches; S slots fill it with other molecules' kernels.

Host code (Python and ctypes) releases the interpreter lock while a slot waits
for its stream, so the slots' launches and host eigensolves overlap each
other's GPU work.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass

import numpy as np
import torch

from .geometry import diatomic_along_z
from .grid import Grid3D
from .physics.molecules import (hcl_hgh_elements, lih_hgh_elements, solve_h2_at_geometry)
from .physics.water import MoleculePoint, solve_rhf_at_geometry
from .solvers.lanczos import default_krylov_dim

KINDS = ("H2", "LiH", "HCl")
GRID_SIZES = (64, 80, 96, 128)
R_EQUILIBRIUM = {"H2": 1.40, "LiH": 3.02, "HCl": 2.41}   # h2.py:22, lih.py:22, hcl.py:28
BOX = {"H2": 10.0, "LiH": 14.0, "HCl": 14.0}              # driver --L defaults
MIXING = {"H2": 0.7, "LiH": 0.5, "HCl": 0.5}              # driver --mixing defaults
MAX_ITERS = {"H2": 30, "LiH": 50, "HCl": 60}              # driver --scf-iters defaults
N_OCC = {"H2": 1, "LiH": 1, "HCl": 4}


@dataclass(frozen=True)
class BatchMolecule:
    """One molecule of a batch: kind, grid points per axis, bond length (bohr)."""

    kind: str
    N: int
    R: float

    @property
    def L(self) -> float:
        return BOX[self.kind]

    @property
    def grid(self) -> Grid3D:
        return Grid3D(self.N, self.L)

    @property
    def n_occ(self) -> int:
        return N_OCC[self.kind]

    @property
    def lanczos_iters(self) -> int:
        m = default_krylov_dim(self.N, self.n_occ)
        return max(m, 200) if (self.kind == "HCl" and self.N >= 80) else m

    def nuclei(self):
        """Along z about the origin: H-H, Li-H (lih.py:61-64), H-Cl (hcl.py:89-99)."""
        if self.kind == "H2":
            return diatomic_along_z(self.R, 1.0)
        if self.kind == "LiH":
            return diatomic_along_z(self.R, 3.0, 1.0)
        if self.kind == "HCl":
            return diatomic_along_z(self.R, 1.0, 17.0)
        raise KeyError(self.kind)

    def cost(self) -> float:
        """Relative work of one SCF iteration (Krylov steps x grid x solves)."""
        return float(self.lanczos_iters) * self.N ** 3 * (1 + self.n_occ)


def draw_uneven_batch(n: int = 64, seed: int = 0) -> list[BatchMolecule]:
    """``n`` molecules: kind, N and R drawn independently from numpy's
    generator seeded with ``seed``."""
    if n < 1:
        raise ValueError("a batch needs at least one molecule")
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, len(KINDS), n)
    sizes = rng.integers(0, len(GRID_SIZES), n)
    scale = rng.uniform(0.85, 1.15, n)
    return [BatchMolecule(KINDS[k], GRID_SIZES[s], float(R_EQUILIBRIUM[KINDS[k]] * f))
            for k, s, f in zip(kinds, sizes, scale)]


def eigensolver_for(mol: BatchMolecule, eigensolver: str, precision: str = "mixed") -> str:
    """``"auto"`` with precision="mixed": scf_rhf's own default, the same as
    the H2O workload's (block Lanczos seeded with the occupied orbitals and
    stopped by its Ritz residuals, on the compressed exchange in every
    iteration; the seeded block holds HCl's degenerate 3pi pair).  With
    another precision: block Lanczos over the whole Krylov dimension for HCl,
    whose 3pi pair is exactly degenerate: on N >= 80 scalar Lanczos with the
    driver's Krylov dimension resolves it into one copy only
    (lanczos.py:83-107, hcl.py:40-51), and on 64^3 it splits the pair by
    rounding (-0.4847 / -0.4833 Ha) and the SCF needs 16 to more than 60
    iterations depending on the bond length, where block Lanczos converges
    every drawn HCl in 14-16; scalar Lanczos for H2 and LiH (no degenerate
    occupied level).  Any other value is used as given."""
    if eigensolver == "auto":
        if precision == "mixed":
            return "auto"
        return "block_lanczos" if mol.kind == "HCl" else "lanczos"
    return eigensolver


def solve_molecule(mol: BatchMolecule, *, max_iters: int | None = None, tol: float = 1e-6,
                   precision: str = "mixed", order: int = 4,
                   eigensolver: str = "lanczos", use_graph: bool | None = None) -> MoleculePoint:
    """The reference driver's single-point RHF for one molecule of a batch, on
    the current stream.  eigensolver="block_lanczos" seeds each iteration with
    the occupied orbitals (for HCl's degenerate 3p pair, which scalar Lanczos
    can resolve into one copy only, lanczos.py:83-107); "auto" picks per
    molecule and precision (``eigensolver_for``)."""
    if max_iters is None:
        max_iters = MAX_ITERS[mol.kind]
    eigensolver = eigensolver_for(mol, eigensolver, precision)
    kw = dict(max_iters=max_iters, tol=tol, mixing=MIXING[mol.kind], order=order,
              lanczos_iters=mol.lanczos_iters, precision=precision, eigensolver=eigensolver,
              use_graph=use_graph)
    nuc = mol.nuclei()
    if mol.kind == "H2":
        return solve_h2_at_geometry(nuc, mol.grid, **kw)
    elements = lih_hgh_elements() if mol.kind == "LiH" else hcl_hgh_elements()
    return solve_rhf_at_geometry(nuc, elements, mol.grid, **kw)


@dataclass
class BatchResult:
    points: list            # MoleculePoint per molecule, in input order
    scf_iterations: int     # sum of every molecule's SCF iterations
    seconds: float          # wall time of the whole batch, setup included
    streams: int

    @property
    def iterations_per_s(self) -> float:
        return self.scf_iterations / self.seconds


def scf_batch(molecules: list[BatchMolecule], *, streams: int = 4, max_iters: int | None = None,
              tol: float = 1e-6, precision: str = "mixed", order: int = 4,
              eigensolver: str = "lanczos") -> BatchResult:
    """Solve every molecule, ``streams`` at a time.

    ``max_iters`` overrides the per-kind driver default for all molecules
    (with ``tol=0`` every molecule then runs exactly that many iterations).
    The wall time runs from the first launch to the end of the last
    molecule's stream."""
    if streams < 1:
        raise ValueError("streams must be >= 1")
    if not torch.cuda.is_available():
        raise RuntimeError("scf_batch needs a ROCm device")
    from . import _lib as L
    L.lib()
    order_idx = sorted(range(len(molecules)), key=lambda i: -molecules[i].cost())
    points: list = [None] * len(molecules)
    errors: list = []
    lock = threading.Lock()
    queue = iter(order_idx)
    dev = torch.device("cuda", torch.cuda.current_device())
    slots = [torch.cuda.Stream(device=dev) for _ in range(min(streams, len(molecules)))]

    def run(slot: torch.cuda.Stream) -> None:
        torch.cuda.set_device(dev)
        with torch.cuda.stream(slot):
            while True:
                with lock:
                    if errors:
                        return
                    i = next(queue, None)
                if i is None:
                    break
                try:
                    points[i] = solve_molecule(molecules[i], max_iters=max_iters, tol=tol,
                                               precision=precision, order=order,
                                               eigensolver=eigensolver, use_graph=use_graph)
                except BaseException as e:  # handed to the caller below
                    with lock:
                        errors.append(e)
                    return
        slot.synchronize()

    # Lanczos graphs only with one slot: HIP refuses some operations (tearing
    # a graph down among them) while another thread's stream records
    use_graph = False if len(slots) > 1 else None
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    # plans evicted mid-batch are released after it: their destructors
    # synchronise the device and would stall the other slots
    with L.deferred_release():
        if len(slots) == 1:
            run(slots[0])
        else:
            threads = [threading.Thread(target=run, args=(s,), daemon=True) for s in slots]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        torch.cuda.synchronize(dev)
        seconds = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return BatchResult(points, int(sum(p.n_iters for p in points)), seconds, len(slots))
