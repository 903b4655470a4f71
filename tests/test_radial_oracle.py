"""The 1D log-radial bare-Coulomb oracle (physics/radial_hydrogen.py), pinned
by the reference's own closed-form tests (test_radial_hydrogen.py,
test_radial_ell.py: E_n = -Z^2/2n^2, l-degeneracy, the 1s->2p radial
dipole 256/(81 sqrt 6)), and used on the GPU to cross-check the eps -> 0
extrapolation of the 3D softened-Coulomb hydrogen atom."""
import math

import numpy as np
import pytest

from oracle import gato_oracle as O


@pytest.mark.parametrize("Z,r_max", [(1.0, 40.0), (2.0, 20.0), (3.0, 15.0)])
def test_hydrogenic_ground_state_energy(Z, r_max):  # test_radial_hydrogen.py:20-36
    E0, _ = O.solve_ground_state(O.LogRadialGrid(400, 0.01, r_max), Z=Z)
    assert abs(E0 + 0.5 * Z * Z) < 1e-4


def test_ground_state_normalised_and_nodeless():  # test_radial_hydrogen.py:39-51
    g = O.LogRadialGrid(400, 0.01, 40.0)
    _, u = O.solve_ground_state(g, 1.0)
    assert abs(O.radial_inner_product(u, u, g) - 1.0) < 1e-10
    u = u if u.sum() > 0 else -u
    assert np.all(u[np.abs(u) > 1e-3] > 0)


def test_grid_convergence_is_monotone():  # test_radial_hydrogen.py:54-70
    errs = [abs(O.solve_ground_state(O.LogRadialGrid(N, 0.01, 40.0), 1.0)[0] + 0.5)
            for N in (100, 200, 400)]
    assert errs[0] >= errs[1] >= errs[2] and errs[-1] < 1e-4


def test_potential_is_unsoftened():  # test_radial_hydrogen.py:73-86
    g = O.LogRadialGrid(300, 0.01, 30.0)
    assert float(np.min(g.r())) < 1e-3
    assert abs(O.solve_ground_state(g, 1.0)[0] + 0.5) < 1e-3


@pytest.mark.parametrize("ell,levels", [
    (1, {2: -0.125, 3: -1.0 / 18, 4: -0.03125}),
    (2, {3: -1.0 / 18, 4: -0.03125, 5: -0.02}),
])
def test_ell_channel_energies(ell, levels):  # test_radial_ell.py:18-43
    E, _ = O.solve_bound_states(O.LogRadialGrid(800, 1e-3, 150.0), 1.0, K=6, ell=ell)
    for n, want in levels.items():
        assert abs(E[n - (ell + 1)] - want) < 1e-4


def test_2s_2p_degenerate_and_dipole():  # test_radial_ell.py:46-98
    g = O.LogRadialGrid(800, 1e-3, 80.0)
    Es, us = O.solve_bound_states(g, 1.0, K=2, ell=0)
    Ep, up = O.solve_bound_states(g, 1.0, K=1, ell=1)
    assert abs(Es[1] - Ep[0]) < 1e-4 and abs(Es[1] + 0.125) < 1e-4
    r = g.r()
    u1s = us[:, 0] * np.sign(us[np.argmin(np.abs(r - 1.0)), 0])
    u2p = up[:, 0] * np.sign(up[np.argmin(np.abs(r - 4.0)), 0])
    d = O.radial_dipole(u2p, u1s, g)
    assert abs(d - 256.0 / (81.0 * math.sqrt(6.0))) < 1e-3
    assert abs(d / math.sqrt(3.0) - 128.0 * math.sqrt(2.0) / 243.0) < 1e-3


@pytest.mark.gpu
def test_gpu_softened_hydrogen_extrapolates_to_radial_bare_coulomb(gpu):
    """3D hydrogen on the GPU (imaginary time, softened -1/sqrt(r^2+eps^2)),
    extrapolated eps -> 0 with the linear+quadratic model, against the bare
    -1/r energy of the independent log-radial oracle: the extrapolation must
    remove most of the softening bias."""
    from gato_amd.convergence import extrapolate_to_zero
    from gato_amd.physics.hydrogen import solve_hydrogen
    E_radial, _ = O.solve_ground_state(O.LogRadialGrid(400, 0.01, 40.0), 1.0)
    N, L = 64, 12.0
    h = L / N
    rows = []
    for f in (0.25, 0.375, 0.5, 0.75, 1.0):
        r = solve_hydrogen(N=N, L=L, epsilon=f * h, n_steps=4000, order=4, verbose=False)
        rows.append((f * h, r.energy, 0, True))
    E0, _ = extrapolate_to_zero(rows)
    E_half = rows[2][1]
    assert abs(E0 - E_radial) < 0.2 * abs(E_half - E_radial)
    assert abs(E0 - E_radial) < 0.01
