"""Writes tests/golden/golden.npz: small fixtures (inputs and expected outputs)
of the FP64 oracle on grids of at most 16^3.

The reference ships no data files (every expected value in its tests is a
closed form or an inline constant), so these fixtures are produced by the
oracle after the oracle itself is pinned by those closed forms
(tests/test_oracle.py).  tests/test_golden.py checks the oracle against them
on the CPU (drift of the restatement) and the HIP path against them on the
GPU (so GPU parity does not depend on re-running the oracle there).

    python tests/golden/make_golden.py      # from the repository root
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import gato_oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")


def cases() -> dict:
    d: dict = {}
    rng = np.random.default_rng(20240611)

    # finite-difference Laplacian, both orders and boundaries (operators.py:41-131)
    N, h = 12, 0.37
    psi = rng.standard_normal((N, N, N))
    d["lap_psi"] = psi
    d["lap_h"] = np.float64(h)
    for order in (2, 4):
        for b in ("dirichlet", "periodic"):
            d[f"lap_{b}_{order}"] = O.laplacian(psi, h, b, order)

    # Hockney Poisson: potential and energy (poisson.py:29-92)
    N, L = 12, 4.8
    g = O.Grid3D(N, L)
    rho = rng.standard_normal((N, N, N))
    d["poi_N"], d["poi_L"] = np.int64(N), np.float64(L)
    d["poi_rho"] = rho
    d["poi_J"] = O.hartree_potential(rho, g)
    d["poi_EH"] = np.float64(O.hartree_energy(rho, g))

    # exchange apply and pair integrals (scf.py:66-85, :223-229)
    orb = rng.standard_normal((N, N, N, 3))
    xpsi = rng.standard_normal((N, N, N))
    d["x_orb"], d["x_psi"] = orb, xpsi
    d["x_Kpsi"] = O.exchange_apply(xpsi, orb, g)
    pairs = []
    for i in range(3):
        for j in range(i, 3):
            dens = orb[..., i] * orb[..., j]
            pairs.append(O.integrate(dens * O.hartree_potential(dens, g), g))
    d["x_pairs"] = np.array(pairs)

    # HGH water: local potential and nonlocal apply (pseudopotentials.py:201-383)
    N, L = 16, 8.0
    g = O.Grid3D(N, L)
    nuc = O.water_initial_geometry(1.81, 104.5)
    el = O.water_hgh_elements()
    d["w_N"], d["w_L"] = np.int64(N), np.float64(L)
    d["w_pos"], d["w_Z"] = nuc.positions, nuc.charges
    d["w_Vloc"] = O.hgh_local_potential(nuc, el, g)
    wpsi = rng.standard_normal((N, N, N))
    d["w_psi"] = wpsi
    d["w_Vnl_psi"] = O.hgh_nonlocal_apply(wpsi, nuc, el, g)

    # Lanczos on the harmonic oscillator (lanczos.py:110-206)
    N, L = 12, 8.0
    g = O.Grid3D(N, L)
    psi0 = rng.standard_normal((N, N, N))
    res = O.lanczos(O.Hamiltonian(g, O.harmonic_oscillator(g), order=4), psi0, 40, 4)
    d["lz_N"], d["lz_L"] = np.int64(N), np.float64(L)
    d["lz_psi0"] = psi0
    d["lz_evals"], d["lz_alphas"], d["lz_betas"] = res.eigenvalues, res.alphas, res.betas

    # helium RHF, 6 iterations from the default start (scf.py:335-444)
    N, L = 16, 8.0
    g = O.Grid3D(N, L)
    he = O.scf_rhf(O.softened_coulomb(g, Z=2.0), g, 1, max_iters=6, tol=1e-14, mixing=0.7,
                   order=4, lanczos_iters=30)
    d["he_N"], d["he_L"] = np.int64(N), np.float64(L)
    d["he_history"] = np.array(he.energy_history)
    d["he_orbital_energies"] = np.array(he.orbital_energies)

    # water RHF energy terms and Hellmann-Feynman forces at fixed orbitals
    # (scf.py:195-265, water.py:181-241)
    N, L = 16, 8.0
    g = O.Grid3D(N, L)
    wres, E, Enn = O.solve_rhf_at_geometry(nuc, el, g, max_iters=4, tol=1e-14, mixing=0.5,
                                           order=4)
    orbs = wres.orbitals
    terms = O.decompose_rhf_energy(orbs, d["w_Vloc"], g, order=4,
                                   V_nl_apply=lambda p: O.hgh_nonlocal_apply(p, nuc, el, g))
    d["we_orbitals"] = orbs
    d["we_terms"] = np.array([terms.kinetic, terms.v_ext, terms.v_nl, terms.hartree,
                              terms.exchange])
    d["we_force"] = O.hellmann_feynman_force(nuc.positions, nuc.charges, orbs, el, g)
    return d


def main():
    d = cases()
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT}: {len(d)} arrays, {os.path.getsize(OUT) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
