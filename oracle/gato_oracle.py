"""FP64 CPU oracle for the real-space RHF hot path (TEST INFRASTRUCTURE ONLY).

This module restates, in numpy/scipy float64, the algorithm of the reference
package (ezequielgarcia/gato, ``src/gato``).  It exists to *check* the HIP
implementation in ``gato_amd``; it is never the thing measured or shipped.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import it.  ``gato_amd`` itself never imports it and has no CPU
fallback.

Every function cites the reference file:line it follows.  Each is pinned by a
test in ``tests/test_oracle_*.py`` against the reference's own known-answer
tests (closed forms, convergence rates, the He SCF constant), against dense
brute force on tiny grids (direct Coulomb sums, dense ``numpy.linalg.eigh``),
or, for the force, against central finite differences of the oracle energy.
Two pieces are "parity unpinned" because only an import of the reference's
third-party dependencies could pin them: the optax Adam trajectory
(``Adam`` below follows Kingma & Ba 2015 with optax's documented defaults) and
jax.random streams (seeded starts here use numpy's own generator).

Layout: fields are C-order (N, N, N) arrays indexed [x][y][z] (z contiguous),
as produced by ``meshgrid(indexing="ij")`` (grid.py:46-49).  Orbital stacks are
(N, N, N, n) with the orbital index last, as in the reference.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Callable

import numpy as np
import scipy.fft as sfft
from scipy.special import erf as _erf

_WORKERS = int(os.environ.get("GATO_ORACLE_WORKERS", os.cpu_count() or 1))


# ---------------------------------------------------------------------------
# Grid (grid.py)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class Grid3D:
    """Cell-centred grid on [-L/2, L/2]^3 (grid.py:23-54)."""

    N: int
    L: float

    @property
    def h(self) -> float:  # grid.py:30-32
        return self.L / self.N

    @property
    def shape(self):
        return (self.N, self.N, self.N)

    @property
    def dV(self) -> float:  # grid.py:38-40
        return self.h ** 3

    def axis(self) -> np.ndarray:  # grid.py:42-44
        return -self.L / 2 + (np.arange(self.N) + 0.5) * self.h

    def coords(self):  # grid.py:46-49
        a = self.axis()
        return np.meshgrid(a, a, a, indexing="ij")

    def radial(self, softening: float = 0.0) -> np.ndarray:  # grid.py:51-54
        X, Y, Z = self.coords()
        return np.sqrt(X * X + Y * Y + Z * Z + softening * softening)


def integrate(psi, grid: Grid3D) -> float:  # grid.py:57-59
    return float(np.sum(psi)) * grid.dV


def inner_product(phi, psi, grid: Grid3D) -> float:  # grid.py:62-64
    return float(np.sum(np.conj(phi) * psi).real) * grid.dV


def norm_sq(psi, grid: Grid3D) -> float:  # grid.py:67-69
    return float(np.sum(np.abs(psi) ** 2)) * grid.dV


def normalize(psi, grid: Grid3D):  # grid.py:72-73
    return psi / math.sqrt(norm_sq(psi, grid))


# ---------------------------------------------------------------------------
# Finite-difference operators (operators.py)
# ---------------------------------------------------------------------------

_O2_NB, _O2_C = 1.0, -2.0                       # operators.py:30-32
_O4_NB1, _O4_NB2, _O4_C = 16.0 / 12.0, -1.0 / 12.0, -30.0 / 12.0  # operators.py:34-37


def laplacian(psi, h: float, boundary: str = "dirichlet", order: int = 2):
    """7-point (order 2) or 13-point (order 4) Laplacian (operators.py:40-119).

    Dirichlet: zero padding by 1 (order 2) or 2 (order 4) cells
    (operators.py:72-79, :99-114); periodic: np.roll (:68-71, :91-98).
    """
    if order not in (2, 4):
        raise ValueError(f"Unknown order {order!r}; expected 2 or 4.")
    if boundary not in ("dirichlet", "periodic"):
        raise ValueError(f"Unknown boundary {boundary!r}; expected 'dirichlet' or 'periodic'.")
    psi = np.asarray(psi, dtype=np.float64)
    if order == 2:
        taps = ((1, _O2_NB),)
        center = 3.0 * _O2_C
    else:
        taps = ((1, _O4_NB1), (2, _O4_NB2))
        center = 3.0 * _O4_C
    acc = center * psi
    if boundary == "periodic":
        for ax in range(3):
            for s, c in taps:
                acc = acc + c * (np.roll(psi, s, axis=ax) + np.roll(psi, -s, axis=ax))
    else:
        R = taps[-1][0]
        p = np.pad(psi, R)
        n = psi.shape
        core = tuple(slice(R, R + n[a]) for a in range(3))
        for ax in range(3):
            for s, c in taps:
                plus = list(core)
                minus = list(core)
                plus[ax] = slice(R + s, R + s + n[ax])
                minus[ax] = slice(R - s, R - s + n[ax])
                acc = acc + c * (p[tuple(plus)] + p[tuple(minus)])
    return acc / (h * h)


def kinetic(psi, h: float, boundary: str = "dirichlet", order: int = 2):
    """T psi = -1/2 lap psi (operators.py:122-130)."""
    return -0.5 * laplacian(psi, h, boundary, order)


@dataclass(frozen=True)
class Hamiltonian:
    """H = T + V (hamiltonian.py:18-35); default order 2 (hamiltonian.py:23)."""

    grid: Grid3D
    V: np.ndarray
    boundary: str = "dirichlet"
    order: int = 2

    def apply(self, psi):
        return kinetic(psi, self.grid.h, self.boundary, self.order) + self.V * psi

    def expectation(self, psi):
        return inner_product(psi, self.apply(psi), self.grid)

    def rayleigh(self, psi):
        return self.expectation(psi) / norm_sq(psi, self.grid)


# ---------------------------------------------------------------------------
# Potentials (potentials.py) and geometry (geometry.py)
# ---------------------------------------------------------------------------

def softened_coulomb(grid: Grid3D, Z: float = 1.0, epsilon=None, center=(0.0, 0.0, 0.0)):
    """-Z / sqrt(|r-r0|^2 + eps^2), eps default h/2 (potentials.py:16-36)."""
    if epsilon is None:
        epsilon = grid.h / 2
    X, Y, Zc = grid.coords()
    x0, y0, z0 = center
    r2 = (X - x0) ** 2 + (Y - y0) ** 2 + (Zc - z0) ** 2
    return -Z / np.sqrt(r2 + epsilon * epsilon)


def multi_center_softened_coulomb(nuclei: "Nuclei", grid: Grid3D, epsilon=None):
    """Sum over nuclei of softened Coulomb wells (potentials.py:39-61)."""
    if epsilon is None:
        epsilon = grid.h / 2
    X, Y, Zc = grid.coords()
    V = np.zeros(grid.shape)
    for R, Zk in zip(nuclei.positions, nuclei.charges):
        r2 = (X - R[0]) ** 2 + (Y - R[1]) ** 2 + (Zc - R[2]) ** 2
        V = V - Zk / np.sqrt(r2 + epsilon * epsilon)
    return V


def harmonic_oscillator(grid: Grid3D, omega: float = 1.0):  # potentials.py:64-71
    X, Y, Z = grid.coords()
    return 0.5 * omega * omega * (X * X + Y * Y + Z * Z)


@dataclass(frozen=True)
class Nuclei:
    """K nuclei: positions (K,3) bohr, charges (K,) (geometry.py:20-28)."""

    positions: np.ndarray
    charges: np.ndarray

    @property
    def n(self) -> int:
        return int(np.asarray(self.charges).shape[0])


def diatomic_along_z(R: float, Z1: float, Z2=None) -> Nuclei:  # geometry.py:31-45
    if Z2 is None:
        Z2 = Z1
    return Nuclei(np.array([[0.0, 0.0, -R / 2], [0.0, 0.0, R / 2]]), np.array([Z1, Z2], dtype=float))


def bond_length(nuclei: Nuclei, i: int, j: int) -> float:  # geometry.py:48-50
    return float(np.linalg.norm(nuclei.positions[i] - nuclei.positions[j]))


def bond_angle(nuclei: Nuclei, i: int, j: int, k: int) -> float:  # geometry.py:53-58
    a = nuclei.positions[i] - nuclei.positions[j]
    b = nuclei.positions[k] - nuclei.positions[j]
    c = np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))
    return float(np.arccos(np.clip(c, -1.0, 1.0)))


def nuclear_repulsion(nuclei: Nuclei) -> float:
    """E_nn = sum_{i<j} Z_i Z_j / R_ij (geometry.py:61-75)."""
    P = np.asarray(nuclei.positions, dtype=float)
    Z = np.asarray(nuclei.charges, dtype=float)
    E = 0.0
    for i in range(len(Z)):
        for j in range(i + 1, len(Z)):
            E += Z[i] * Z[j] / np.linalg.norm(P[i] - P[j])
    return float(E)


def nuclear_repulsion_grad(nuclei: Nuclei) -> np.ndarray:
    """dE_nn/dR, analytic derivative of ``nuclear_repulsion``."""
    P = np.asarray(nuclei.positions, dtype=float)
    Z = np.asarray(nuclei.charges, dtype=float)
    g = np.zeros_like(P)
    for i in range(len(Z)):
        for j in range(i + 1, len(Z)):
            d = P[i] - P[j]
            r = np.linalg.norm(d)
            gi = -Z[i] * Z[j] * d / r ** 3
            g[i] += gi
            g[j] -= gi
    return g


def center_of_mass(nuclei: Nuclei) -> np.ndarray:  # geometry.py:78-83
    Z = np.asarray(nuclei.charges, dtype=float)
    return (Z[:, None] * nuclei.positions).sum(axis=0) / Z.sum()


def recenter(nuclei: Nuclei) -> Nuclei:  # geometry.py:86-91
    return Nuclei(nuclei.positions - center_of_mass(nuclei), nuclei.charges)


# ---------------------------------------------------------------------------
# HGH pseudopotentials (pseudopotentials.py)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class HGHProjector:  # pseudopotentials.py:81-96
    ell: int
    r: float
    h: tuple

    @property
    def n(self) -> int:
        return len(self.h)


@dataclass(frozen=True)
class HGHParams:  # pseudopotentials.py:99-112
    symbol: str
    Z_ion: float
    r_loc: float
    c: tuple
    projectors: tuple = ()


# Parameter tables (pseudopotentials.py:122-185): HGH 1998 table I values.
HGH_TABLE = {
    "H": HGHParams("H", 1.0, 0.2, (-4.0663326, 0.6778322, 0.0, 0.0), ()),
    "O": HGHParams("O", 6.0, 0.24762086, (-16.5803180, 2.39570952, 0.0, 0.0),
                   (HGHProjector(0, 0.22178614, ((18.26691718,),)),)),
    "Li": HGHParams("Li", 1.0, 0.78755305, (-1.89261247, 0.28605968, 0.0, 0.0),
                    (HGHProjector(0, 0.66637518, ((1.85881111,),)),
                     HGHProjector(1, 1.07930561, ((-0.00589504,),)))),
    "Cl": HGHParams("Cl", 7.0, 0.41, (-6.86475431, 0.0, 0.0, 0.0),
                    (HGHProjector(0, 0.33820832, ((9.06223968, -1.96193036),
                                                  (-1.96193036, 5.06568240))),
                     HGHProjector(1, 0.37613709, ((4.46587640,),)))),
}


def lookup(symbol: str) -> HGHParams:  # pseudopotentials.py:188-194
    if symbol not in HGH_TABLE:
        raise KeyError(f"No HGH parameters for {symbol!r}.")
    return HGH_TABLE[symbol]


def total_valence_charge(elements) -> float:  # pseudopotentials.py:256-258
    return float(sum(p.Z_ion for p in elements))


def _grid_pts(grid: Grid3D):
    X, Y, Z = grid.coords()
    return np.stack([X, Y, Z], axis=-1)


def _hgh_local_one_center(grid_pts, R, p: HGHParams):
    """V_loc for one nucleus (pseudopotentials.py:201-227)."""
    diff = grid_pts - np.asarray(R)
    r2 = np.sum(diff * diff, axis=-1)
    r = np.sqrt(r2 + 1e-30)
    rho = r / p.r_loc
    erf_term = -p.Z_ion * _erf(r / (math.sqrt(2.0) * p.r_loc)) / r
    c1, c2, c3, c4 = p.c
    rho2 = rho * rho
    poly = c1 + c2 * rho2 + c3 * rho2 * rho2 + c4 * rho2 * rho2 * rho2
    return erf_term + np.exp(-0.5 * rho2) * poly


def hgh_local_potential(nuclei: Nuclei, elements, grid: Grid3D):
    """Sum of local parts over nuclei (pseudopotentials.py:230-253)."""
    if len(elements) != nuclei.n:
        raise ValueError(f"HGH parameter list length {len(elements)} != number of nuclei {nuclei.n}")
    gp = _grid_pts(grid)
    V = np.zeros(grid.shape)
    for k, p in enumerate(elements):
        V = V + _hgh_local_one_center(gp, nuclei.positions[k], p)
    return V


def _hgh_local_radial_derivative(r, p: HGHParams):
    """dV_loc/dr for one centre, analytic derivative of _hgh_local_one_center."""
    a = math.sqrt(2.0) * p.r_loc
    e = _erf(r / a)
    de = (2.0 / math.sqrt(math.pi)) * np.exp(-(r / a) ** 2) / a
    d_erf = -p.Z_ion * (de / r - e / (r * r))
    rho2 = (r / p.r_loc) ** 2
    c1, c2, c3, c4 = p.c
    poly = c1 + c2 * rho2 + c3 * rho2 ** 2 + c4 * rho2 ** 3
    dpoly_drho2 = c2 + 2 * c3 * rho2 + 3 * c4 * rho2 ** 2
    drho2_dr = 2.0 * r / p.r_loc ** 2
    g = np.exp(-0.5 * rho2)
    d_poly = g * (-0.5 * drho2_dr * poly + dpoly_drho2 * drho2_dr)
    return d_erf + d_poly


def projector_radial(r, ell: int, i_idx: int, r_ell: float):
    """p_{l i}(r) (pseudopotentials.py:265-277), i_idx 1-based."""
    n = ell + (4 * i_idx - 1) / 2.0
    norm = math.sqrt(2.0) / (r_ell ** n * math.sqrt(math.gamma(n)))
    power = ell + 2 * (i_idx - 1)
    return norm * (r ** power) * np.exp(-(r * r) / (2.0 * r_ell * r_ell))


def angular_factor(ell: int, m: int, hat):
    """Real Y_lm (pseudopotentials.py:280-309); l=1 order m=-1:y, 0:z, 1:x."""
    if ell == 0 and m == 0:
        return np.full(hat.shape[:-1], 1.0 / math.sqrt(4.0 * math.pi))
    if ell == 1:
        c = math.sqrt(3.0 / (4.0 * math.pi))
        return c * hat[..., {-1: 1, 0: 2, 1: 0}[m]]
    if ell == 2:
        x, y, z = hat[..., 0], hat[..., 1], hat[..., 2]
        if m == -2:
            return math.sqrt(15.0 / (4.0 * math.pi)) * x * y
        if m == -1:
            return math.sqrt(15.0 / (4.0 * math.pi)) * y * z
        if m == 0:
            return 0.25 * math.sqrt(5.0 / math.pi) * (3.0 * z * z - 1.0)
        if m == 1:
            return math.sqrt(15.0 / (4.0 * math.pi)) * x * z
        if m == 2:
            return 0.25 * math.sqrt(15.0 / math.pi) * (x * x - y * y)
    raise NotImplementedError(f"Real Y_{{{ell},{m}}} not implemented (need l <= 2).")


def projector_field(grid_pts, R, ell, m, i_idx, r_ell):
    """p_{l i}(r) Y_lm(r^) centred on R (pseudopotentials.py:312-327)."""
    diff = grid_pts - np.asarray(R)
    r = np.sqrt(np.sum(diff * diff, axis=-1) + 1e-30)
    hat = diff / r[..., None]
    return projector_radial(r, ell, i_idx, r_ell) * angular_factor(ell, m, hat)


def _projector_list(nuclei: Nuclei, elements):
    """Enumerate (k, l, m, n, r_l, h) channels in the reference's loop order
    (pseudopotentials.py:338-356, :377-382)."""
    out = []
    for k, p in enumerate(elements):
        for proj in p.projectors:
            for m in range(-proj.ell, proj.ell + 1):
                out.append((k, proj.ell, m, proj.n, proj.r, np.asarray(proj.h, dtype=float)))
    return out


def hgh_nonlocal_apply(psi, nuclei: Nuclei, elements, grid: Grid3D):
    """sum_k sum_lm sum_ij p_i h_ij <p_j|psi> on the full grid
    (pseudopotentials.py:330-383)."""
    gp = _grid_pts(grid)
    out = np.zeros_like(psi, dtype=float)
    for (k, ell, m, n, r_l, h) in _projector_list(nuclei, elements):
        ps = [projector_field(gp, nuclei.positions[k], ell, m, i + 1, r_l) for i in range(n)]
        ov = np.array([inner_product(pj, psi, grid) for pj in ps])
        coeffs = h @ ov
        for i in range(n):
            out = out + coeffs[i] * ps[i]
    return out


def hgh_apply(psi, nuclei, elements, grid, V_local=None):  # pseudopotentials.py:390-405
    if V_local is None:
        V_local = hgh_local_potential(nuclei, elements, grid)
    return V_local * psi + hgh_nonlocal_apply(psi, nuclei, elements, grid)


# ---------------------------------------------------------------------------
# Hockney free-space Poisson (solvers/poisson.py)
# ---------------------------------------------------------------------------

def coulomb_kernel(N: int, h: float, epsilon: float):
    """Softened 1/r on the (2N)^3 wrap-around grid (poisson.py:29-44)."""
    M = 2 * N
    idx = np.arange(M)
    disp = np.where(idx < N, idx * h, (idx - M) * h)
    X = disp[:, None, None]
    Y = disp[None, :, None]
    Z = disp[None, None, :]
    return 1.0 / np.sqrt(X * X + Y * Y + Z * Z + epsilon * epsilon)


_GHAT_CACHE: dict = {}


def _green_hat(N: int, h: float, epsilon: float):
    key = (N, float(h), float(epsilon))
    g = _GHAT_CACHE.get(key)
    if g is None:
        g = sfft.rfftn(coulomb_kernel(N, h, epsilon), workers=_WORKERS)
        if len(_GHAT_CACHE) > 4:
            _GHAT_CACHE.clear()
        _GHAT_CACHE[key] = g
    return g


def hartree_potential(rho, grid: Grid3D, epsilon=None):
    """J = irfftn(rfftn(pad(rho)) * rfftn(G))[:N,:N,:N] * dV (poisson.py:47-79).

    The transform of G is cached per (N, h, eps); the reference rebuilds it on
    every call (poisson.py:51,54) but the values are identical.
    """
    if epsilon is None:
        epsilon = grid.h / 2
    N = grid.N
    M = 2 * N
    rho = np.asarray(rho, dtype=np.float64)
    R = sfft.rfftn(rho, s=(M, M, M), workers=_WORKERS)  # zero-pads to (M,M,M)
    J = sfft.irfftn(R * _green_hat(N, grid.h, epsilon), s=(M, M, M), workers=_WORKERS)
    return J[:N, :N, :N] * grid.dV


def hartree_energy(rho, grid: Grid3D, epsilon=None) -> float:
    """E_H = 1/2 int rho J (poisson.py:82-92)."""
    return 0.5 * integrate(rho * hartree_potential(rho, grid, epsilon), grid)


# ---------------------------------------------------------------------------
# Lanczos (solvers/lanczos.py)
# ---------------------------------------------------------------------------

@dataclass
class LanczosResult:  # lanczos.py:41-46
    eigenvalues: np.ndarray
    eigenstates: np.ndarray
    n_iters: int
    converged: bool
    alphas: np.ndarray = field(default=None, repr=False)
    betas: np.ndarray = field(default=None, repr=False)


def default_krylov_dim(N: int, n_eigenstates: int = 1) -> int:
    """max(24 n + 24, floor(5N/4)) (lanczos.py:49-107)."""
    return max(24 * n_eigenstates + 24, (5 * N) // 4)


def lanczos(H, psi0, n_iters: int = 80, n_eigenstates: int = 4, tol: float = 1e-10) -> LanczosResult:
    """Single-vector Lanczos with full MGS reorthogonalisation
    (lanczos.py:110-206)."""
    grid = H.grid
    shape = psi0.shape
    v0 = psi0 / math.sqrt(norm_sq(psi0, grid))
    V = np.zeros((n_iters + 1,) + shape)
    V[0] = v0
    alphas = np.zeros(n_iters)
    betas = np.zeros(n_iters)
    for j in range(n_iters):
        vj = V[j]
        w = H.apply(vj)
        if j > 0:                                   # lanczos.py:146-148
            w = w - betas[j - 1] * V[j - 1]
        alpha = inner_product(vj, w, grid)          # lanczos.py:150
        alphas[j] = alpha
        w = w - alpha * vj
        for k in range(j + 1):                      # lanczos.py:155-159 (MGS)
            w = w - inner_product(V[k], w, grid) * V[k]
        beta = math.sqrt(norm_sq(w, grid))
        betas[j] = beta
        V[j + 1] = w / beta if beta > 0 else 0.0    # lanczos.py:163-167
    T = np.diag(alphas)                             # lanczos.py:176-179
    if n_iters >= 2:
        off = betas[: n_iters - 1]
        T = T + np.diag(off, 1) + np.diag(off, -1)
    evals, evecs = np.linalg.eigh(T)                # lanczos.py:180
    K = min(n_eigenstates, n_iters)
    # Eigenvector signs are a gauge the reference leaves to LAPACK; the SCF
    # seeds its next Lanczos run with sum_i phi_i (scf.py:417), so the sign
    # steers unconverged iterations.  Fix it: largest-magnitude component > 0
    # (first index on ties), the rule the HIP path's eigensolver applies.
    for k in range(K):
        i = int(np.argmax(np.abs(evecs[:, k])))
        if evecs[i, k] < 0:
            evecs[:, k] = -evecs[:, k]
    psi = np.tensordot(evecs[:, :K].T, V[:n_iters], axes=(1, 0))   # lanczos.py:193-194
    converged = bool(np.any(betas[: max(n_iters - 1, 1)] < tol))  # lanczos.py:199
    return LanczosResult(evals[:K], np.moveaxis(psi, 0, -1), int(n_iters), converged, alphas, betas)


def _cholqr(W, dV):
    """Rows of W as vectors: W = R^T Q with Q orthonormal (dV Q Q^T = I) and R
    upper triangular with a positive diagonal, by Cholesky QR applied twice."""
    G = dV * (W @ W.T)
    R1 = np.linalg.cholesky(G).T
    Q = np.linalg.solve(R1.T, W)
    G2 = dV * (Q @ Q.T)
    R2 = np.linalg.cholesky(G2).T
    Q = np.linalg.solve(R2.T, Q)
    return Q, R2 @ R1


def block_lanczos(H, X0, n_steps: int, n_eigenstates: int, tol: float = 1e-10) -> LanczosResult:
    """Block Lanczos with full reorthogonalisation, seeded with the b columns
    of X0 (N, N, N, b).  The reference has none: its Lanczos docstring
    (lanczos.py:83-107) names it as the fix for degenerate levels.  This is
    the recurrence of lanczos.py:140-194 with vectors replaced by blocks:
    W = H Q_i - Q_{i-1} B_{i-1}^T, A_i = Q_i^T W, W -= Q_i A_i, two classical
    Gram-Schmidt sweeps against the whole basis, W = Q_{i+1} B_i by Cholesky
    QR; T is block tridiagonal with A_i on the diagonal and B_i below it.
    Ritz vectors carry the scalar path's sign gauge.  ``alphas``/``betas``
    hold the A_i / B_i blocks."""
    grid = H.grid
    dV = grid.dV
    shape = X0.shape[:3]
    b = X0.shape[3]
    P = int(np.prod(shape))
    Q, _ = _cholqr(np.moveaxis(X0, -1, 0).reshape(b, P), dV)
    blocks = [Q]
    A, B = [], []
    Qprev, Bprev = None, None
    for i in range(n_steps):
        W = np.stack([H.apply(q.reshape(shape)).ravel() for q in Q])
        if i > 0:
            W = W - Bprev @ Qprev
        Ai = dV * (Q @ W.T)
        W = W - Ai.T @ Q
        Vall = np.concatenate(blocks)
        for _ in range(2):
            W = W - (dV * (Vall @ W.T)).T @ Vall
        A.append(0.5 * (Ai + Ai.T))
        if i + 1 < n_steps:
            Qn, Bi = _cholqr(W, dV)
        else:  # the last B_i is not part of T
            Qn, Bi = None, np.linalg.cholesky(dV * (W @ W.T)).T
        B.append(Bi)
        Qprev, Bprev, Q = Q, Bi, Qn
        if Qn is not None:
            blocks.append(Qn)
    m = n_steps * b
    T = np.zeros((m, m))
    for i in range(n_steps):
        T[i * b:(i + 1) * b, i * b:(i + 1) * b] = A[i]
        if i + 1 < n_steps:
            T[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = B[i]
            T[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = B[i].T
    evals, evecs = np.linalg.eigh(T)
    K = min(n_eigenstates, m)
    for k in range(K):
        j = int(np.argmax(np.abs(evecs[:, k])))
        if evecs[j, k] < 0:
            evecs[:, k] = -evecs[:, k]
    V = np.concatenate(blocks)[:m]
    psi = (evecs[:, :K].T @ V).reshape((K,) + tuple(shape))
    diagB = np.abs(np.array([np.diag(Bi) for Bi in B[: max(n_steps - 1, 1)]]))
    converged = bool(np.any(diagB < tol))
    return LanczosResult(evals[:K], np.moveaxis(psi, 0, -1), m, converged, np.array(A),
                         np.array(B))


# ---------------------------------------------------------------------------
# RHF (scf.py)
# ---------------------------------------------------------------------------

def density(orbitals):
    """rho = 2 sum_i |phi_i|^2 (scf.py:61-63)."""
    return 2.0 * np.sum(np.abs(orbitals) ** 2, axis=-1)


def exchange_apply(psi, orbitals, grid: Grid3D, epsilon=None):
    """K psi = sum_j phi_j J[phi_j* psi] (scf.py:66-85)."""
    out = np.zeros_like(psi, dtype=float)
    for j in range(orbitals.shape[-1]):
        phi = orbitals[..., j]
        out = out + phi * hartree_potential(np.conj(phi) * psi, grid, epsilon)
    return out


@dataclass(frozen=True)
class RHFFock:
    """F psi = T psi + V_ext psi + V_nl psi + V_H psi - K psi (scf.py:88-158)."""

    grid: Grid3D
    V_ext: np.ndarray
    orbitals: np.ndarray
    V_H: np.ndarray = None
    boundary: str = "dirichlet"
    order: int = 4
    epsilon: float = None
    V_nl_apply: Callable = None

    def density(self):
        return density(self.orbitals)

    def hartree(self):
        if self.V_H is not None:
            return self.V_H
        return hartree_potential(self.density(), self.grid, self.epsilon)

    def apply(self, psi):
        eps = self.grid.h / 2 if self.epsilon is None else self.epsilon
        h_psi = kinetic(psi, self.grid.h, self.boundary, self.order) + self.V_ext * psi
        if self.V_nl_apply is not None:
            h_psi = h_psi + self.V_nl_apply(psi)
        return h_psi + self.hartree() * psi - exchange_apply(psi, self.orbitals, self.grid, eps)


@dataclass(frozen=True)
class RHFEnergyTerms:  # scf.py:165-191
    kinetic: float
    v_ext: float
    v_nl: float
    hartree: float
    exchange: float

    @property
    def one_body(self):
        return self.kinetic + self.v_ext + self.v_nl

    @property
    def total(self):
        return self.kinetic + self.v_ext + self.v_nl + self.hartree + self.exchange


def decompose_rhf_energy(orbitals, V_ext, grid: Grid3D, *, boundary="dirichlet", order=4,
                         epsilon=None, V_nl_apply=None) -> RHFEnergyTerms:
    """Energy decomposition (scf.py:194-265)."""
    if epsilon is None:
        epsilon = grid.h / 2
    n = orbitals.shape[-1]
    T = Ve = Vnl = 0.0
    for i in range(n):
        phi = orbitals[..., i]
        T += inner_product(phi, kinetic(phi, grid.h, boundary, order), grid)
        Ve += inner_product(phi, V_ext * phi, grid)
        if V_nl_apply is not None:
            Vnl += inner_product(phi, V_nl_apply(phi), grid)
    EH = hartree_energy(density(orbitals), grid, epsilon)
    Ex = 0.0
    for i in range(n):
        phi = orbitals[..., i]
        Ex -= inner_product(phi, exchange_apply(phi, orbitals, grid, epsilon), grid)
    return RHFEnergyTerms(2 * T, 2 * Ve, 2 * Vnl, EH, Ex)


def rhf_energy(orbitals, V_ext, grid, boundary="dirichlet", order=4, epsilon=None, V_nl_apply=None):
    return decompose_rhf_energy(orbitals, V_ext, grid, boundary=boundary, order=order,
                                epsilon=epsilon, V_nl_apply=V_nl_apply).total  # scf.py:268-286


def default_initial_guess_vector(grid: Grid3D):
    """Core-H Lanczos starter (scf.py:310-320)."""
    X, Y, Z = grid.coords()
    r = np.sqrt(X * X + Y * Y + Z * Z + 1e-6)
    core = np.exp(-4.0 * r)
    valence = np.exp(-r)
    pert = 0.1 * (X + Y + Z) + 0.02 * (X * Y + Y * Z + X * Z)
    return (core + valence) * (1.0 + pert)


def default_initial_orbitals(V_ext, grid, n_occ, boundary, order, lanczos_iters):
    """Core-Hamiltonian guess (scf.py:289-322)."""
    H = Hamiltonian(grid, V_ext, boundary, order)
    return lanczos(H, default_initial_guess_vector(grid), lanczos_iters, n_occ).eigenstates


@dataclass
class SCFResult:  # scf.py:325-332
    orbitals: np.ndarray
    energy: float
    orbital_energies: np.ndarray
    n_iters: int
    converged: bool
    energy_history: list


def scf_rhf(V_ext, grid: Grid3D, n_occ: int, *, initial_orbitals=None, max_iters=50, tol=1e-6,
            mixing=1.0, boundary="dirichlet", order=4, lanczos_iters=None, epsilon=None,
            V_nl_apply=None, max_lanczos_steps_hook=None) -> SCFResult:
    """Closed-shell SCF with linear density mixing (scf.py:335-444)."""
    if lanczos_iters is None:
        lanczos_iters = default_krylov_dim(grid.N, n_occ)
    if initial_orbitals is None:
        orbitals = default_initial_orbitals(V_ext, grid, n_occ, boundary, order, lanczos_iters)
    else:
        orbitals = np.asarray(initial_orbitals, dtype=float)
    history = []
    rho = density(orbitals)
    E_prev = E = float("inf")
    eps_orb = np.zeros(n_occ)
    converged = False
    it = 0
    for it in range(1, max_iters + 1):
        V_H = hartree_potential(rho, grid, epsilon)                  # scf.py:399
        fock = RHFFock(grid, V_ext, orbitals, V_H, boundary, order, epsilon, V_nl_apply)
        seed = np.sum(orbitals, axis=-1)                             # scf.py:417
        res = lanczos(fock, seed, lanczos_iters, n_occ)              # scf.py:418
        new_orb = res.eigenstates
        eps_orb = res.eigenvalues
        rho = mixing * density(new_orb) + (1.0 - mixing) * rho       # scf.py:422-423
        orbitals = new_orb
        E = float(rhf_energy(orbitals, V_ext, grid, boundary, order, epsilon, V_nl_apply))
        history.append(E)
        if abs(E - E_prev) < tol:                                    # scf.py:432-435
            converged = True
            break
        E_prev = E
    return SCFResult(orbitals, E, eps_orb, it, converged, history)


# ---------------------------------------------------------------------------
# Molecule drivers, BO energy and Hellmann-Feynman force (physics/water.py)
# ---------------------------------------------------------------------------

def water_initial_geometry(R_OH: float = 1.85, angle_deg: float = 100.0) -> Nuclei:
    """O at origin, H in the xz-plane at +-half angle (water.py:66-82)."""
    half = math.radians(angle_deg) / 2.0
    h1 = [R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]
    h2 = [-R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]
    return Nuclei(np.array([[0.0, 0.0, 0.0], h1, h2]), np.array([8.0, 1.0, 1.0]))


def water_hgh_elements():  # water.py:85-87
    return (lookup("O"), lookup("H"), lookup("H"))


def pp_nuclear_repulsion(nuclei: Nuclei, elements) -> float:
    """E_nn with valence charges Z_ion (water.py:100-110)."""
    if nuclei.n < 2:
        return 0.0
    return nuclear_repulsion(Nuclei(nuclei.positions, np.array([p.Z_ion for p in elements])))


def solve_rhf_at_geometry(nuclei: Nuclei, elements, grid: Grid3D, *, max_iters=60, tol=1e-6,
                          mixing=0.5, order=4, lanczos_iters=None, initial_orbitals=None):
    """RHF with HGH PP at fixed nuclei (water.py:129-174).  Returns
    (SCFResult, E_total, E_nn)."""
    V_local = hgh_local_potential(nuclei, elements, grid)
    n_el = int(total_valence_charge(elements))
    if n_el % 2 != 0:
        raise ValueError(f"Closed-shell RHF requires even electron count; got {n_el}")
    V_nl = (lambda psi: hgh_nonlocal_apply(psi, nuclei, elements, grid))
    if not any(p.projectors for p in elements):
        V_nl = None
    res = scf_rhf(V_local, grid, n_el // 2, initial_orbitals=initial_orbitals, max_iters=max_iters,
                  tol=tol, mixing=mixing, order=order, lanczos_iters=lanczos_iters, V_nl_apply=V_nl)
    E_nn = pp_nuclear_repulsion(nuclei, elements)
    return res, res.energy + E_nn, E_nn


def bo_energy(positions, charges, orbitals, elements, grid: Grid3D, order: int = 4) -> float:
    """E(R) at fixed orbitals = decompose_rhf_energy + E_nn(PP) (water.py:181-215)."""
    nuc = Nuclei(np.asarray(positions, dtype=float), np.asarray(charges, dtype=float))
    V_local = hgh_local_potential(nuc, elements, grid)
    V_nl = (lambda psi: hgh_nonlocal_apply(psi, nuc, elements, grid))
    terms = decompose_rhf_energy(orbitals, V_local, grid, boundary="dirichlet", order=order,
                                 epsilon=None, V_nl_apply=V_nl)
    return terms.total + pp_nuclear_repulsion(nuc, elements)


def _projector_grad_fields(gp, R, ell, m, i_idx, r_ell):
    """Gradient of p(r - R) with respect to R (= -grad_r p), shape (...,3).

    Derived from p = norm * d^power * exp(-d^2/(2 r^2)) * Y_lm(d_vec/d).
    For l = 0: grad_r p = f'(d) d_vec / d * Y00.  For l = 1, p = c * g(d) *
    d_a with g = f/d, so grad_r p = c (g'(d) d_vec/d d_a + g e_a).
    """
    diff = gp - np.asarray(R)
    r = np.sqrt(np.sum(diff * diff, axis=-1) + 1e-30)
    n = ell + (4 * i_idx - 1) / 2.0
    norm = math.sqrt(2.0) / (r_ell ** n * math.sqrt(math.gamma(n)))
    power = ell + 2 * (i_idx - 1)
    ex = np.exp(-(r * r) / (2.0 * r_ell * r_ell))
    if ell == 0:
        # f = norm r^p e; f' = norm e (p r^{p-1} - r^{p+1}/r_l^2)
        fp = norm * ex * (power * r ** (power - 1) if power > 0 else 0.0) - norm * ex * r ** (power + 1) / r_ell ** 2
        Y = 1.0 / math.sqrt(4.0 * math.pi)
        grad_r = (fp * Y / r)[..., None] * diff
    elif ell == 1:
        c = math.sqrt(3.0 / (4.0 * math.pi))
        a = {-1: 1, 0: 2, 1: 0}[m]
        q = power - 1  # g = norm r^q e
        g = norm * r ** q * ex
        gp_ = norm * ex * (q * r ** (q - 1) if q > 0 else 0.0) - norm * ex * r ** (q + 1) / r_ell ** 2
        grad_r = c * (gp_ / r * diff[..., a])[..., None] * diff
        grad_r[..., a] += c * g
    else:
        raise NotImplementedError("projector gradients implemented for l <= 1 (all tabulated elements)")
    return -grad_r


def hellmann_feynman_force(positions, charges, orbitals, elements, grid: Grid3D) -> np.ndarray:
    """F = -dE/dR at fixed orbitals (water.py:221-241), analytic.

    dE/dR_k = int rho dV_loc^k/dR_k + 4 sum_i sum_{lm,ab} h_ab <phi_i|dp_a/dR><p_b|phi_i>
              + dE_nn/dR_k.
    The reference evaluates the same derivative with jax.grad; this oracle
    states it in closed form and is pinned against central finite differences
    of ``bo_energy``.
    """
    P = np.asarray(positions, dtype=float)
    gp = _grid_pts(grid)
    rho = density(orbitals)
    grad = np.zeros_like(P)
    for k, p in enumerate(elements):
        diff = gp - P[k]
        r = np.sqrt(np.sum(diff * diff, axis=-1) + 1e-30)
        dV = _hgh_local_radial_derivative(r, p)
        # d/dR of V(|r-R|) = -V'(d) (r-R)/d
        grad[k] += -np.einsum("xyz,xyzc->c", rho * dV / r, diff) * grid.dV
    nuc = Nuclei(P, np.asarray(charges, dtype=float))
    for (k, ell, m, n, r_l, h) in _projector_list(nuc, elements):
        ps = [projector_field(gp, P[k], ell, m, i + 1, r_l) for i in range(n)]
        dps = [_projector_grad_fields(gp, P[k], ell, m, i + 1, r_l) for i in range(n)]
        for o in range(orbitals.shape[-1]):
            phi = orbitals[..., o]
            ov = np.array([np.sum(pj * phi) * grid.dV for pj in ps])
            dov = np.array([np.einsum("xyz,xyzc->c", phi, dp) * grid.dV for dp in dps])  # (n,3)
            # 2 (occupation) * d/dR [sum_ab h_ab ov_a ov_b] = 2 * 2 * sum_ab h_ab dov_a ov_b
            grad[k] += 4.0 * np.einsum("ab,ac,b->c", h, dov, ov)
    Zion = np.array([p.Z_ion for p in elements])
    if len(elements) >= 2:
        grad += nuclear_repulsion_grad(Nuclei(P, Zion))
    return -grad


class Adam:
    """Textbook Adam (Kingma & Ba 2015) with optax.adam's documented defaults
    b1=0.9, b2=0.999, eps=1e-8, eps_root=0 (geometry_opt.py:62-71 uses
    optax.adam(lr)).  Parity unpinned: no reference test holds a trajectory."""

    def __init__(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.m = None
        self.v = None
        self.t = 0

    def update(self, g):
        g = np.asarray(g, dtype=float)
        if self.m is None:
            self.m = np.zeros_like(g)
            self.v = np.zeros_like(g)
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return -self.lr * mh / (np.sqrt(vh) + self.eps)


def optimize_geometry_alternating(initial_nuclei: Nuclei, *, solve_at_geometry, bo_grad, geom_steps,
                                  learning_rate=0.05, on_step=None, recenter_charges=None):
    """Alternating SCF + Adam + recenter (geometry_opt.py:28-78)."""
    nuclei = recenter(initial_nuclei)
    opt = Adam(learning_rate)
    current = solve_at_geometry(nuclei, None)
    traj = [current]
    for step in range(geom_steps):
        g = bo_grad(current)
        upd = opt.update(g)
        new_pos = current.nuclei.positions + upd
        nuclei = recenter(Nuclei(new_pos, current.nuclei.charges))
        current = solve_at_geometry(nuclei, current)
        if on_step is not None:
            on_step(step, current)
        traj.append(current)
    return current, traj


# ---------------------------------------------------------------------------
# Imaginary time (solvers/imag_time.py) and H2+ (physics/h2_plus.py)
# ---------------------------------------------------------------------------

def init_lcao(nuclei: Nuclei, grid: Grid3D, softening=None):
    """sum_k Z_k^{3/2} exp(-Z_k sqrt(|r-R_k|^2 + s^2)), s default h/2
    (ansatz/grid.py:53-76)."""
    if softening is None:
        softening = grid.h / 2
    X, Y, Zc = grid.coords()
    psi = np.zeros(grid.shape)
    for R, Zk in zip(nuclei.positions, nuclei.charges):
        r = np.sqrt((X - R[0]) ** 2 + (Y - R[1]) ** 2 + (Zc - R[2]) ** 2 + softening ** 2)
        psi = psi + Zk ** 1.5 * np.exp(-Zk * r)
    return psi


@dataclass
class ImagTimeHistory:  # solvers/imag_time.py:24-27
    steps: list
    energies: list


def imaginary_time(H: Hamiltonian, psi0, dt=None, n_steps=2000, log_every=50):
    """Euler steps psi <- psi - dt H psi, renormalised; dt default h^2/10
    (solvers/imag_time.py:30-65)."""
    if dt is None:
        dt = H.grid.h ** 2 / 10.0
    psi = psi0 / np.sqrt(norm_sq(psi0, H.grid))
    steps, energies = [], []
    for i in range(n_steps):
        psi_new = psi - dt * H.apply(psi)
        psi = psi_new / np.sqrt(norm_sq(psi_new, H.grid))
        if (i % log_every == 0) or (i == n_steps - 1):
            steps.append(i)
            energies.append(float(H.rayleigh(psi)))
    return psi, ImagTimeHistory(steps, energies)


def h2p_bo_energy(positions, charges, psi, grid: Grid3D, epsilon: float, order: int = 4) -> float:
    """<psi|T + V(R)|psi>/<psi|psi> + E_nn (physics/h2_plus.py:66-88)."""
    nuc = Nuclei(np.asarray(positions, dtype=float), np.asarray(charges, dtype=float))
    V = multi_center_softened_coulomb(nuc, grid, epsilon)
    Hpsi = kinetic(psi, grid.h, "dirichlet", order) + V * psi
    return inner_product(psi, Hpsi, grid) / norm_sq(psi, grid) + nuclear_repulsion(nuc)


def h2p_hellmann_feynman_force(positions, charges, psi, grid: Grid3D, epsilon: float,
                               order: int = 4) -> np.ndarray:
    """F = -dE/dR at fixed psi (h2_plus.py:99-108; jax.grad replaced by the
    analytic derivative of the softened Coulomb wells):
    dV/dR_a = -Z (r - R)_a / (|r-R|^2 + eps^2)^{3/2}."""
    P = np.asarray(positions, dtype=float)
    Zs = np.asarray(charges, dtype=float)
    rho = psi * psi / norm_sq(psi, grid)
    X, Y, Zc = grid.coords()
    F = np.zeros_like(P)
    for k in range(len(Zs)):
        d = (X - P[k, 0], Y - P[k, 1], Zc - P[k, 2])
        q = (d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + epsilon ** 2) ** 1.5
        for a in range(3):
            dV = -Zs[k] * d[a] / q
            F[k, a] = -integrate(rho * dV, grid)
    return F - nuclear_repulsion_grad(Nuclei(P, Zs))


@dataclass
class SolvePoint:  # h2_plus.py:115-121
    nuclei: Nuclei
    psi: np.ndarray
    energy: float
    electronic_energy: float
    nuclear_repulsion: float


def h2p_solve_electronic(nuclei: Nuclei, grid: Grid3D, epsilon=None, n_steps=2000,
                         order=4) -> SolvePoint:
    """Imaginary time from the LCAO guess (h2_plus.py:124-146)."""
    if epsilon is None:
        epsilon = grid.h / 2
    V = multi_center_softened_coulomb(nuclei, grid, epsilon)
    H = Hamiltonian(grid, V, "dirichlet", order)
    psi0 = normalize(init_lcao(nuclei, grid), grid)
    psi, _ = imaginary_time(H, psi0, n_steps=n_steps, log_every=n_steps + 1)
    E_nn = nuclear_repulsion(nuclei)
    E_el = float(H.rayleigh(psi))
    return SolvePoint(nuclei, psi, E_el + E_nn, E_el, E_nn)


def h2p_bo_curve(R_values, grid: Grid3D, Z=1.0, epsilon=None, n_steps=2000, order=4):
    """h2_plus.py:153-168."""
    return [h2p_solve_electronic(diatomic_along_z(float(R), Z1=Z), grid, epsilon, n_steps, order)
            for R in R_values]


def kinetic_energy(psi, grid: Grid3D, boundary="dirichlet", order=2) -> float:
    """<T> / <psi|psi> (observables.py:15-21)."""
    return inner_product(psi, kinetic(psi, grid.h, boundary, order), grid) / norm_sq(psi, grid)


def potential_energy(psi, V, grid: Grid3D) -> float:  # observables.py:24-27
    return inner_product(psi, V * psi, grid) / norm_sq(psi, grid)


def virial_ratio(psi, V, grid: Grid3D, boundary="dirichlet", order=2) -> float:
    """2<T>/|<V>| (observables.py:30-42)."""
    return 2.0 * kinetic_energy(psi, grid, boundary, order) / abs(potential_energy(psi, V, grid))


@dataclass
class HydrogenResult:  # physics/hydrogen.py:36-46
    psi: np.ndarray
    energy: float
    kinetic: float
    potential: float
    virial_ratio: float
    history: list


def solve_hydrogen(N=48, L=10.0, Z=1.0, epsilon=None, n_steps=3000, order=2) -> HydrogenResult:
    """Imaginary time from exp(-Z r_soft), r softened by h/2 (physics/hydrogen.py:49-96,
    ansatz/grid.py:37-39)."""
    grid = Grid3D(N, L)
    V = softened_coulomb(grid, Z=Z, epsilon=epsilon)
    H = Hamiltonian(grid, V, "dirichlet", order)
    psi0 = np.exp(-Z * grid.radial(softening=grid.h / 2))
    psi, hist = imaginary_time(H, psi0, n_steps=n_steps, log_every=max(1, n_steps // 20))
    psi = normalize(psi, grid)
    T = kinetic_energy(psi, grid, order=order)
    Vx = potential_energy(psi, V, grid)
    return HydrogenResult(psi, T + Vx, T, Vx, virial_ratio(psi, V, grid, order=order),
                          list(zip(hist.steps, hist.energies)))


# ---------------------------------------------------------------------------
# 1D log-radial solver (physics/radial_hydrogen.py:74-280): a bare -Z/r
# dense-eigh oracle, independent of the 3D softened-Coulomb discretisation
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class LogRadialGrid:
    """Cell-centred xi in (0, 1), r = r_min (exp(alpha xi) - 1),
    alpha = log(r_max/r_min + 1) (radial_hydrogen.py:74-99)."""
    N: int
    r_min: float
    r_max: float

    @property
    def alpha(self) -> float:
        return math.log(self.r_max / self.r_min + 1.0)

    @property
    def h_xi(self) -> float:
        return 1.0 / self.N

    def xi(self):
        return (np.arange(self.N) + 0.5) * self.h_xi

    def r(self):
        return self.r_min * (np.exp(self.alpha * self.xi()) - 1.0)

    def r_prime(self):
        return self.alpha * (self.r() + self.r_min)

    def r_double_prime(self):
        return self.alpha * self.r_prime()


def _radial_pad(u, left_sign):
    """Signed reflection through r = 0 (parity (-1)^(l+1)), zeros at xi = 1
    (radial_hydrogen.py:102-125)."""
    return np.concatenate([left_sign * u[:2][::-1], u, np.zeros(2)])


def radial_laplacian(u, grid: LogRadialGrid, ell: int = 0):
    """u''(r) = u''_xi / r'^2 - r''/r'^3 u'_xi with 4th-order xi stencils
    (radial_hydrogen.py:128-160)."""
    h = grid.h_xi
    p = _radial_pad(u, -1 if ell % 2 == 0 else 1)
    du = (p[:-4] - 8 * p[1:-3] + 8 * p[3:-1] - p[4:]) / (12 * h)
    d2u = (-p[:-4] + 16 * p[1:-3] - 30 * p[2:-2] + 16 * p[3:-1] - p[4:]) / (12 * h * h)
    rp, rpp = grid.r_prime(), grid.r_double_prime()
    return d2u / (rp * rp) - rpp / rp ** 3 * du


def radial_hamiltonian(u, grid: LogRadialGrid, Z=1.0, ell=0):
    """-1/2 u'' - Z/r u + l(l+1)/(2 r^2) u (radial_hydrogen.py:163-190)."""
    r = grid.r()
    V = -Z / r + (0.5 * ell * (ell + 1) / (r * r) if ell else 0.0)
    return -0.5 * radial_laplacian(u, grid, ell) + V * u


def radial_inner_product(u, v, grid: LogRadialGrid) -> float:  # radial_hydrogen.py:193-197
    return float(grid.h_xi * np.sum(u * v * grid.r_prime()))


def solve_bound_states(grid: LogRadialGrid, Z=1.0, K=5, ell=0):
    """Lowest K states: dense H, symmetrised by W^{1/2} H W^{-1/2} with
    W = diag(r'), numpy eigh, mapped back and W-normalised
    (radial_hydrogen.py:200-255)."""
    N = grid.N
    H = np.zeros((N, N))
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1.0
        H[:, i] = radial_hamiltonian(e, grid, Z, ell)
    ws = np.sqrt(grid.r_prime())
    Hs = (ws[:, None] * H) / ws[None, :]
    Hs = 0.5 * (Hs + Hs.T)
    w, U = np.linalg.eigh(Hs)
    states = U[:, :K] / ws[:, None]
    states = states / np.sqrt(grid.h_xi * np.sum(states * states * grid.r_prime()[:, None], axis=0))
    return w[:K], states


def solve_ground_state(grid: LogRadialGrid, Z=1.0):  # radial_hydrogen.py:207-219
    E, u = solve_bound_states(grid, Z, K=1, ell=0)
    return float(E[0]), u[:, 0]


def radial_dipole(u_final, u_initial, grid: LogRadialGrid) -> float:
    """<u_f| r |u_i> (radial_hydrogen.py:258-280)."""
    return float(grid.h_xi * np.sum(u_final * grid.r() * u_initial * grid.r_prime()))
