"""Accuracy studies on the GPU path: helium E(eps) -> eps = 0 softening
extrapolation and grid-spacing / box-size / geometry convergence of the HGH
molecules (reference: benchmarks/helium_softening_extrapolation.py:45-95,
benchmarks/grid_convergence.py:30-265).

Every SCF runs through ``scf_rhf`` / ``solve_rhf_at_geometry`` on the HIP
kernels; the fits are small host least-squares problems.
"""
from __future__ import annotations

import math
import time

import numpy as np

from .geometry import bond_angle, bond_length
from .grid import Grid3D
from .physics import molecules as mol
from .physics import water as water_mod
from .potentials import softened_coulomb
from .scf import scf_rhf

HE_E_RHF_REF = -2.8617  # published closed-shell He RHF (helium_softening_extrapolation.py:36)
HE_E_EXACT = -2.9037


# --------------------------------------------------------------------------- softening


def default_epsilons(h: float) -> list:
    """0.25h .. 1.0h around the default eps = h/2 (helium_softening_extrapolation.py:160-163)."""
    return [0.25 * h, 0.375 * h, 0.5 * h, 0.75 * h, 1.0 * h]


def softening_sweep(N: int, L: float, epsilons, mixing: float = 0.7, max_iters: int = 40,
                    order: int = 4, lanczos_iters: int = 40, precision: str = "fp64",
                    tol: float = 1e-6, verbose: bool = False) -> list:
    """One He RHF SCF per eps at fixed (N, L), with the same eps in the
    nuclear well and the Hartree/exchange kernel
    (helium_softening_extrapolation.py:45-76).  Rows (eps, E, n_iters, converged)."""
    grid = Grid3D(N=N, L=L)
    rows = []
    for eps in epsilons:
        V = softened_coulomb(grid, Z=2.0, epsilon=eps)
        res = scf_rhf(V, grid, 1, max_iters=max_iters, tol=tol, mixing=mixing, order=order,
                      lanczos_iters=lanczos_iters, epsilon=eps, precision=precision)
        rows.append((float(eps), float(res.energy), int(res.n_iters), bool(res.converged)))
        if verbose:
            print(f"  eps = {eps:.5f}   E = {res.energy:+.5f} Ha   iters={res.n_iters:2d} "
                  f"{'converged' if res.converged else 'NOT converged'}", flush=True)
    return rows


def extrapolate_to_zero(rows, model: str = "linear_quadratic"):
    """Least-squares fit of E(eps) = E0 + a eps (+ b eps^2); returns (E0, coef)
    (helium_softening_extrapolation.py:79-98)."""
    eps = np.array([r[0] for r in rows], dtype=float)
    E = np.array([r[1] for r in rows], dtype=float)
    if model == "quadratic":
        A = np.column_stack([np.ones_like(eps), eps ** 2])
    elif model == "linear":
        A = np.column_stack([np.ones_like(eps), eps])
    elif model == "linear_quadratic":
        A = np.column_stack([np.ones_like(eps), eps, eps ** 2])
    else:
        raise ValueError(f"unknown model {model!r}")
    coef, *_ = np.linalg.lstsq(A, E, rcond=None)
    return float(coef[0]), coef


# --------------------------------------------------------------------------- power law


def fit_power_law(xs, ys):
    """y(x) = y_inf + C x^p by a scan over p in [0.5, 6] (1101 points), each a
    2x2 linear least squares; returns (y_inf, C, p, rms) or NaNs for < 3
    points (grid_convergence.py:117-142)."""
    xs = np.asarray(xs, dtype=float)
    ys = np.asarray(ys, dtype=float)
    if len(xs) < 3:
        return float("nan"), float("nan"), float("nan"), float("nan")
    best = None
    for p in np.linspace(0.5, 6.0, 1101):
        A = np.column_stack([np.ones_like(xs), xs ** p])
        coef, *_ = np.linalg.lstsq(A, ys, rcond=None)
        rms = float(np.sqrt(np.mean((ys - A @ coef) ** 2)))
        if best is None or rms < best[3]:
            best = (float(coef[0]), float(coef[1]), float(p), rms)
    return best


# --------------------------------------------------------------------------- molecules


class WaterCase:  # grid_convergence.py:50-78
    name = "H2O"
    default_L = 14.0
    R0 = 1.81
    angle0 = 104.5

    def nuclei(self):
        return water_mod.water_initial_geometry(R_OH=self.R0, angle_deg=self.angle0)

    def elements(self):
        return water_mod.water_hgh_elements()

    def solve(self, nuclei, grid, **kw):
        return water_mod.solve_rhf_at_geometry(nuclei, self.elements(), grid, **kw)

    def optimize(self, grid, **kw):
        res = water_mod.optimize_water_geometry(self.nuclei(), self.elements(), grid,
                                                verbose=False, **kw)
        n = res.final.nuclei
        return {"E": res.final.energy, "R": float(bond_length(n, 0, 1)),
                "angle": math.degrees(float(bond_angle(n, 1, 0, 2))),
                "converged": res.final.converged}


class HClCase:  # grid_convergence.py:81-106
    name = "HCl"
    default_L = 14.0
    R0 = 2.41

    def nuclei(self):
        return mol.hcl_initial_geometry(R_HCl=self.R0)

    def elements(self):
        return mol.hcl_hgh_elements()

    def solve(self, nuclei, grid, **kw):
        return water_mod.solve_rhf_at_geometry(nuclei, self.elements(), grid, **kw)

    def optimize(self, grid, **kw):
        res = mol.optimize_diatomic_geometry(self.nuclei(), self.elements(), grid,
                                             verbose=False, **kw)
        n = res.final.nuclei
        return {"E": res.final.energy, "R": float(bond_length(n, 0, 1)), "angle": float("nan"),
                "converged": res.final.converged}


CASES = {"water": WaterCase, "hcl": HClCase}


def sweep_energy(case, Ns, L: float, order: int = 4, scf_iters: int = 60, verbose: bool = False,
                 **solve_kw) -> dict:
    """E vs h at fixed L and geometry; unconverged points are left out of the
    fit (grid_convergence.py:156-176)."""
    rows, used = [], []
    nuclei = case.nuclei()
    for N in Ns:
        grid = Grid3D(N=N, L=L)
        t0 = time.perf_counter()
        pt = case.solve(nuclei, grid, max_iters=scf_iters, order=order, **solve_kw)
        row = {"N": N, "h": grid.h, "E": pt.energy, "n_iters": pt.n_iters,
               "converged": pt.converged, "seconds": time.perf_counter() - t0}
        rows.append(row)
        if pt.converged:
            used.append((grid.h, pt.energy))
        if verbose:
            print(f"{N:5d} {grid.h:9.5f} {pt.energy:+15.8f} {pt.n_iters:7d} {pt.converged}",
                  flush=True)
    fit = fit_power_law([u[0] for u in used], [u[1] for u in used])
    return {"rows": rows, "fit": fit}


def sweep_box(case, Ls, target_h: float, order: int = 4, scf_iters: int = 60,
              verbose: bool = False, **solve_kw) -> dict:
    """E vs L at (as nearly as an even N allows) fixed h; warns when h drifts,
    since box and resolution error are then confounded; fits against 1/L
    (grid_convergence.py:179-236)."""
    rows, used = [], []
    nuclei = case.nuclei()
    for L in Ls:
        N = int(round(L / target_h / 2)) * 2
        grid = Grid3D(N=N, L=L)
        pt = case.solve(nuclei, grid, max_iters=scf_iters, order=order, **solve_kw)
        rows.append({"L": L, "N": N, "h": grid.h, "E": pt.energy, "converged": pt.converged})
        if pt.converged:
            used.append((L, pt.energy, grid.h))
        if verbose:
            print(f"{L:6.1f} {N:5d} {grid.h:9.5f} {pt.energy:+15.8f} {pt.converged}", flush=True)
    hs = [u[2] for u in used]
    drift = bool(hs) and (max(hs) - min(hs)) > 1e-4
    if drift and verbose:
        print(f"  WARNING: h drifted over {min(hs):.5f}..{max(hs):.5f}; box and resolution "
              "error are confounded")
    fit = fit_power_law([1.0 / u[0] for u in used], [u[1] for u in used]) if len(used) >= 3 \
        else (float("nan"),) * 4
    return {"rows": rows, "fit": fit, "h_drift": drift}


def sweep_geometry(case, Ns, L: float, order: int = 4, geom_steps: int = 30, lr: float = 0.05,
                   scf_iters: int = 60, verbose: bool = False) -> dict:
    """Relaxed geometry vs h (grid_convergence.py:239-265): R, angle and E each
    fitted to h -> 0 over the converged points."""
    rows, used = [], []
    for N in Ns:
        grid = Grid3D(N=N, L=L)
        out = case.optimize(grid, geom_steps=geom_steps, learning_rate=lr,
                            scf_max_iters=scf_iters, order=order)
        rows.append(dict(out, N=N, h=grid.h))
        if out["converged"]:
            used.append((grid.h, out))
        if verbose:
            print(f"{N:5d} {grid.h:9.5f} R = {out['R']:.5f} angle = {out['angle']:.3f} "
                  f"E = {out['E']:+.8f} {out['converged']}", flush=True)
    hs = [u[0] for u in used]
    fits = {"R": fit_power_law(hs, [u[1]["R"] for u in used]),
            "E": fit_power_law(hs, [u[1]["E"] for u in used])}
    if used and not math.isnan(used[0][1]["angle"]):
        fits["angle"] = fit_power_law(hs, [u[1]["angle"] for u in used])
    return {"rows": rows, "fits": fits}
