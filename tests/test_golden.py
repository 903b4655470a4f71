"""Committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the FP64 oracle).

CPU: the oracle reproduces every fixture (guards the restatement against
drift).  GPU: the HIP path matches every fixture to the same tolerances as
the live-oracle parity tests, without running the oracle on the GPU host."""
import os

import numpy as np
import pytest

from oracle import gato_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


@pytest.fixture(scope="module")
def gd():
    with np.load(GOLDEN, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) /
                 max(float(np.max(np.abs(b))), 1e-300))


# ----------------------------------------------------------------------------- CPU


def test_oracle_stencils_reproduce_golden(gd):
    for order in (2, 4):
        for b in ("dirichlet", "periodic"):
            got = O.laplacian(gd["lap_psi"], float(gd["lap_h"]), b, order)
            assert relerr(got, gd[f"lap_{b}_{order}"]) < 1e-14


def test_oracle_poisson_and_exchange_reproduce_golden(gd):
    g = O.Grid3D(int(gd["poi_N"]), float(gd["poi_L"]))
    assert relerr(O.hartree_potential(gd["poi_rho"], g), gd["poi_J"]) < 1e-13
    assert abs(O.hartree_energy(gd["poi_rho"], g) - float(gd["poi_EH"])) < 1e-12 * abs(float(gd["poi_EH"]))
    assert relerr(O.exchange_apply(gd["x_psi"], gd["x_orb"], g), gd["x_Kpsi"]) < 1e-13


def test_oracle_hgh_and_lanczos_reproduce_golden(gd):
    g = O.Grid3D(int(gd["w_N"]), float(gd["w_L"]))
    nuc = O.Nuclei(gd["w_pos"], gd["w_Z"])
    el = O.water_hgh_elements()
    assert relerr(O.hgh_local_potential(nuc, el, g), gd["w_Vloc"]) < 1e-14
    assert relerr(O.hgh_nonlocal_apply(gd["w_psi"], nuc, el, g), gd["w_Vnl_psi"]) < 1e-13
    gl = O.Grid3D(int(gd["lz_N"]), float(gd["lz_L"]))
    res = O.lanczos(O.Hamiltonian(gl, O.harmonic_oscillator(gl), order=4), gd["lz_psi0"], 40, 4)
    assert np.max(np.abs(res.eigenvalues - gd["lz_evals"])) < 1e-11


def test_oracle_scf_and_forces_reproduce_golden(gd):
    g = O.Grid3D(int(gd["he_N"]), float(gd["he_L"]))
    he = O.scf_rhf(O.softened_coulomb(g, Z=2.0), g, 1, max_iters=6, tol=1e-14, mixing=0.7,
                   order=4, lanczos_iters=30)
    assert np.max(np.abs(np.array(he.energy_history) - gd["he_history"])) < 1e-10
    gw = O.Grid3D(int(gd["w_N"]), float(gd["w_L"]))
    nuc = O.Nuclei(gd["w_pos"], gd["w_Z"])
    el = O.water_hgh_elements()
    F = O.hellmann_feynman_force(nuc.positions, nuc.charges, gd["we_orbitals"], el, gw)
    assert np.max(np.abs(F - gd["we_force"])) < 1e-12


# ----------------------------------------------------------------------------- GPU


def _dev(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


def _host(t):
    return t.detach().double().cpu().numpy()


@pytest.mark.gpu
def test_hip_stencils_match_golden(gpu, gd):
    import torch
    import gato_amd as G
    for order in (2, 4):
        for b in ("dirichlet", "periodic"):
            ref = gd[f"lap_{b}_{order}"]
            got = _host(G.laplacian(_dev(gd["lap_psi"]), float(gd["lap_h"]), b, order))
            assert relerr(got, ref) < 1e-13
            got32 = _host(G.laplacian(_dev(gd["lap_psi"], torch.float32), float(gd["lap_h"]), b,
                                      order))
            assert relerr(got32, ref) < 5e-6


@pytest.mark.gpu
def test_hip_poisson_and_exchange_match_golden(gpu, gd):
    import gato_amd as G
    from gato_amd.solvers.poisson import PoissonPlan
    gg = G.Grid3D(int(gd["poi_N"]), float(gd["poi_L"]))
    assert relerr(_host(G.hartree_potential(_dev(gd["poi_rho"]), gg)), gd["poi_J"]) < 1e-11
    EH = float(G.hartree_energy(_dev(gd["poi_rho"]), gg))
    assert abs(EH - float(gd["poi_EH"])) < 1e-10 * abs(float(gd["poi_EH"]))
    assert relerr(_host(G.exchange_apply(_dev(gd["x_psi"]), _dev(gd["x_orb"]), gg)),
                  gd["x_Kpsi"]) < 1e-11
    plan = PoissonPlan(gg.N, gg.h, gg.h / 2, _dev(gd["x_psi"]).dtype)
    pairs = _host(plan.pair_integrals(_dev(np.moveaxis(gd["x_orb"], -1, 0))))
    assert np.max(np.abs(pairs - gd["x_pairs"]) / np.abs(gd["x_pairs"])) < 1e-10


@pytest.mark.gpu
def test_hip_hgh_and_lanczos_match_golden(gpu, gd):
    import gato_amd as G
    from gato_amd.physics.water import water_hgh_elements
    gg = G.Grid3D(int(gd["w_N"]), float(gd["w_L"]))
    nuc = G.make_nuclei(gd["w_pos"], gd["w_Z"])
    el = water_hgh_elements()
    assert relerr(_host(G.hgh_local_potential(nuc, el, gg)), gd["w_Vloc"]) < 1e-13
    assert relerr(_host(G.hgh_nonlocal_apply(_dev(gd["w_psi"]), nuc, el, gg)),
                  gd["w_Vnl_psi"]) < 1e-11
    gl = G.Grid3D(int(gd["lz_N"]), float(gd["lz_L"]))
    res = G.lanczos(G.Hamiltonian(gl, G.harmonic_oscillator(gl), order=4), _dev(gd["lz_psi0"]),
                    40, 4)
    assert np.max(np.abs(res.eigenvalues.numpy() - gd["lz_evals"])) < 1e-9
    assert np.max(np.abs(res.alphas - gd["lz_alphas"])) < 1e-6
    assert np.max(np.abs(res.betas - gd["lz_betas"])) < 1e-6


@pytest.mark.gpu
def test_hip_scf_energies_and_forces_match_golden(gpu, gd):
    import torch
    import gato_amd as G
    from gato_amd.physics.water import water_hgh_elements
    from gato_amd.pseudopotentials import nonlocal_plan
    gh = G.Grid3D(int(gd["he_N"]), float(gd["he_L"]))
    he = G.scf_rhf(G.softened_coulomb(gh, Z=2.0), gh, 1, max_iters=6, tol=1e-14, mixing=0.7,
                   order=4, lanczos_iters=30)
    assert np.max(np.abs(np.array(he.energy_history) - gd["he_history"])) < 1e-8
    gw = G.Grid3D(int(gd["w_N"]), float(gd["w_L"]))
    nuc = G.make_nuclei(gd["w_pos"], gd["w_Z"])
    el = water_hgh_elements()
    V = G.hgh_local_potential(nuc, el, gw)
    terms = G.decompose_rhf_energy(_dev(gd["we_orbitals"]), V, gw, order=4,
                                   V_nl_apply=nonlocal_plan(nuc, el, gw, torch.float64))
    got = np.array([float(terms.kinetic), float(terms.v_ext), float(terms.v_nl),
                    float(terms.hartree), float(terms.exchange)])
    assert np.max(np.abs(got - gd["we_terms"])) < 1e-9
    F = G.hellmann_feynman_force(torch.tensor(gd["w_pos"]), torch.tensor(gd["w_Z"]),
                                 _dev(gd["we_orbitals"]), el, gw)
    assert np.max(np.abs(F.numpy() - gd["we_force"])) < 1e-8
