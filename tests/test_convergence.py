"""Softening extrapolation and grid-convergence studies
(benchmarks/helium_softening_extrapolation.py, benchmarks/grid_convergence.py).

CPU: the fits recover exactly known parameters from synthetic data.  GPU: the
sweeps' SCF energies equal the oracle's at the same (N, L, eps), and the
eps -> 0 extrapolation moves the helium energy towards the published RHF
value."""
import numpy as np
import pytest

from oracle import gato_oracle as O


def test_extrapolate_to_zero_recovers_intercept():
    from gato_amd.convergence import extrapolate_to_zero
    eps = [0.05, 0.075, 0.1, 0.15, 0.2]
    for model, f in [("linear", lambda e: -2.9 + 1.3 * e),
                     ("quadratic", lambda e: -2.9 + 4.0 * e * e),
                     ("linear_quadratic", lambda e: -2.9 + 1.3 * e - 2.5 * e * e)]:
        E0, coef = extrapolate_to_zero([(e, f(e), 10, True) for e in eps], model)
        assert abs(E0 + 2.9) < 1e-12
    with pytest.raises(ValueError):
        extrapolate_to_zero([(0.1, 0.0, 1, True)], "cubic")


def test_fit_power_law_recovers_order():
    from gato_amd.convergence import fit_power_law
    hs = np.array([0.3, 0.25, 0.2, 0.175, 0.15])
    y = -17.0 + 1.6 * hs ** 2.3
    y_inf, C, p, rms = fit_power_law(hs, y)
    assert abs(p - 2.3) < 1e-9 and abs(y_inf + 17.0) < 1e-10 and abs(C - 1.6) < 1e-9
    assert rms < 1e-12
    assert all(np.isnan(v) for v in fit_power_law([0.1, 0.2], [1.0, 2.0]))


def test_default_epsilons():
    from gato_amd.convergence import default_epsilons
    assert default_epsilons(0.2) == pytest.approx([0.05, 0.075, 0.1, 0.15, 0.2])


@pytest.mark.gpu
def test_softening_sweep_matches_oracle(gpu):
    from gato_amd.convergence import softening_sweep
    N, L = 20, 8.0
    g = O.Grid3D(N, L)
    eps_list = [0.1, 0.2, 0.3]
    rows = softening_sweep(N, L, eps_list, max_iters=4, tol=1e-14, lanczos_iters=40)
    for (eps, E, n_it, conv) in rows:
        ref = O.scf_rhf(O.softened_coulomb(g, Z=2.0, epsilon=eps), g, 1, max_iters=4, tol=1e-14,
                        mixing=0.7, order=4, lanczos_iters=40, epsilon=eps)
        assert n_it == ref.n_iters == 4
        assert abs(E - ref.energy) < 1e-8


@pytest.mark.gpu
def test_sweep_energy_matches_oracle(gpu):
    from gato_amd.convergence import WaterCase, sweep_energy
    out = sweep_energy(WaterCase(), [16, 20], 8.0, order=4, scf_iters=3, tol=1e-14)
    nuc = O.water_initial_geometry(1.81, 104.5)
    for row in out["rows"]:
        res, E, Enn = O.solve_rhf_at_geometry(nuc, O.water_hgh_elements(), O.Grid3D(row["N"], 8.0),
                                              max_iters=3, tol=1e-14, mixing=0.5, order=4)
        assert abs(row["E"] - E) < 1e-8
    assert np.isnan(out["fit"][2])  # unconverged points are left out of the fit


@pytest.mark.gpu
def test_helium_softening_extrapolation_reduces_bias(gpu):
    """helium_softening_extrapolation.py: E(eps) rises with eps, and the
    eps -> 0 fit lands closer to the published RHF energy than E(h/2)."""
    from gato_amd.convergence import (HE_E_RHF_REF, default_epsilons, extrapolate_to_zero,
                                      softening_sweep)
    N, L = 48, 10.0
    rows = softening_sweep(N, L, default_epsilons(L / N), max_iters=40, lanczos_iters=40)
    assert all(r[3] for r in rows)
    E = [r[1] for r in rows]
    assert all(a < b for a, b in zip(E, E[1:]))
    E0, _ = extrapolate_to_zero(rows)
    assert abs(E0 - HE_E_RHF_REF) < abs(E[2] - HE_E_RHF_REF)
