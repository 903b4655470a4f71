"""GPU parity of imaginary-time propagation, the LCAO guess, H2+ and the
hydrogen driver against the FP64 oracle, plus the reference's own H2+ and
hydrogen tests (test_h2_plus.py, test_hydrogen.py) run on the HIP path."""
import numpy as np
import pytest
import torch

from oracle import gato_oracle as O

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def host(t):
    return t.detach().double().cpu().numpy()


def test_imaginary_time_matches_oracle(gpu):
    import gato_amd as G
    g, gg = O.Grid3D(16, 8.0), G.Grid3D(16, 8.0)
    psi0 = 1.0 + 0.2 * np.random.default_rng(11).standard_normal(g.shape)
    ref, rh = O.imaginary_time(O.Hamiltonian(g, O.harmonic_oscillator(g), order=4), psi0,
                               n_steps=300, log_every=50)
    H = G.Hamiltonian(gg, G.harmonic_oscillator(gg), order=4)
    psi, hist = G.imaginary_time(H, dev(psi0), n_steps=300, log_every=50)
    assert hist.steps == rh.steps
    assert np.max(np.abs(np.array(hist.energies) - np.array(rh.energies))) < 1e-11
    assert np.max(np.abs(host(psi) - ref)) < 1e-10
    psi32, h32 = G.imaginary_time(H, dev(psi0, torch.float32), n_steps=300, log_every=50)
    assert psi32.dtype == torch.float32
    assert np.max(np.abs(np.array(h32.energies) - np.array(rh.energies))) < 2e-5


@pytest.mark.parametrize("n_steps,log_every", [(0, 5), (1, 50), (7, 3), (12, 1)])
def test_imaginary_time_logging_intervals_match_oracle(gpu, n_steps, log_every):
    """Each logging interval is one library call ending normalised: the
    history, the state and the n_steps = 0 case follow the reference loop."""
    import gato_amd as G
    g, gg = O.Grid3D(16, 8.0), G.Grid3D(16, 8.0)
    psi0 = 1.0 + 0.2 * np.random.default_rng(5).standard_normal(g.shape)
    ref, rh = O.imaginary_time(O.Hamiltonian(g, O.harmonic_oscillator(g), order=2), psi0,
                               n_steps=n_steps, log_every=log_every)
    H = G.Hamiltonian(gg, G.harmonic_oscillator(gg), order=2)
    psi, hist = G.imaginary_time(H, dev(psi0), n_steps=n_steps, log_every=log_every)
    assert hist.steps == rh.steps
    assert np.max(np.abs(np.array(hist.energies) - np.array(rh.energies)), initial=0.0) < 1e-12
    assert np.max(np.abs(host(psi) - ref)) < 1e-12


def test_init_lcao_matches_oracle(gpu):
    import gato_amd as G
    g, gg = O.Grid3D(20, 9.0), G.Grid3D(20, 9.0)
    nuc_o = O.Nuclei(np.array([[0.1, -0.3, 0.4], [0.0, 0.2, -1.3]]), np.array([1.0, 3.0]))
    nuc = G.make_nuclei(nuc_o.positions, nuc_o.charges)
    ref = O.init_lcao(nuc_o, g)
    got = host(G.init_lcao(nuc, gg))
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(ref))


def test_h2p_solve_energy_and_force_match_oracle(gpu):
    import gato_amd as G
    from gato_amd.physics import h2_plus as HP
    g, gg = O.Grid3D(24, 10.0), G.Grid3D(24, 10.0)
    ref = O.h2p_solve_electronic(O.diatomic_along_z(2.2, 1.0), g, n_steps=400)
    pt = HP.solve_electronic(G.diatomic_along_z(2.2, 1.0), gg, n_steps=400)
    assert abs(pt.energy - ref.energy) < 1e-10
    assert abs(pt.nuclear_repulsion - ref.nuclear_repulsion) < 1e-14
    eps = g.h / 2
    F_ref = O.h2p_hellmann_feynman_force(ref.nuclei.positions, ref.nuclei.charges, ref.psi, g, eps)
    F = HP.hellmann_feynman_force(pt.nuclei.positions, pt.nuclei.charges, pt.psi, gg, eps)
    assert np.max(np.abs(F.numpy() - F_ref)) < 1e-9
    E = float(HP.bo_energy(pt.nuclei.positions, pt.nuclei.charges, pt.psi, gg, eps))
    assert abs(E - O.h2p_bo_energy(ref.nuclei.positions, ref.nuclei.charges, ref.psi, g, eps)) < 1e-10


def test_h2p_hellmann_feynman_matches_finite_difference(gpu):
    """test_h2_plus.py:54-76 on the HIP path."""
    import gato_amd as G
    from gato_amd.physics import h2_plus as HP
    grid = G.Grid3D(32, 10.0)
    point = HP.solve_electronic(G.diatomic_along_z(2.2, 1.0), grid, n_steps=1500, order=4)
    eps = grid.h / 2
    F = HP.hellmann_feynman_force(point.nuclei.positions, point.nuclei.charges, point.psi, grid,
                                  eps, order=4)
    d = 1e-3
    plus = point.nuclei.positions.clone()
    minus = point.nuclei.positions.clone()
    plus[1, 2] += d
    minus[1, 2] -= d
    Ep = float(HP.bo_energy(plus, point.nuclei.charges, point.psi, grid, eps, 4))
    Em = float(HP.bo_energy(minus, point.nuclei.charges, point.psi, grid, eps, 4))
    fd = -(Ep - Em) / (2 * d)
    assert float(F[1, 2]) == pytest.approx(fd, rel=1e-3, abs=1e-4)


def test_h2p_bo_curve_has_bound_minimum(gpu):
    """test_h2_plus.py:32-51 on the HIP path."""
    import gato_amd as G
    from gato_amd.physics import h2_plus as HP
    Rs = [1.0, 1.5, 2.0, 2.5, 3.0, 4.0, 6.0]
    curve = HP.bo_curve(Rs, grid=G.Grid3D(40, 10.0), n_steps=1500, order=4)
    E = [p.energy for p in curve]
    assert E[-1] < -0.40
    i = int(np.argmin(E))
    assert 1.5 <= Rs[i] <= 2.5
    assert E[i] < E[-1] - 0.03


def test_h2p_geometry_optimization_recovers_bond_length(gpu):
    """test_h2_plus.py:79-101 on the HIP path."""
    import gato_amd as G
    from gato_amd.physics import h2_plus as HP
    res = HP.optimize_geometry(G.diatomic_along_z(2.6, 1.0), G.Grid3D(32, 10.0),
                               electronic_steps=1200, geom_steps=20, learning_rate=0.05, order=4)
    R = float(G.bond_length(res.final.nuclei, 0, 1))
    assert 1.7 <= R <= 2.3
    assert abs(R - 2.6) > 0.2
    assert res.final.energy < -0.5
    assert len(res.trajectory) == 20


def test_hydrogen_matches_oracle(gpu):
    from gato_amd.physics.hydrogen import solve_hydrogen
    ref = O.solve_hydrogen(N=24, L=10.0, n_steps=600, order=4)
    res = solve_hydrogen(N=24, L=10.0, n_steps=600, order=4, verbose=False)
    assert abs(res.energy - ref.energy) < 1e-10
    assert abs(res.virial_ratio - ref.virial_ratio) < 1e-10
    assert [s for s, _ in res.history] == [s for s, _ in ref.history]


def test_hydrogen_reference_bounds(gpu):
    """test_hydrogen.py:13-45 on the HIP path."""
    from gato_amd.physics.hydrogen import solve_hydrogen
    r = solve_hydrogen(N=40, L=10.0, n_steps=1500, verbose=False)
    assert -0.55 < r.energy < -0.44
    assert abs(r.virial_ratio - 1.0) < 0.08
    for Z, L in [(1.0, 10.0), (2.0, 6.0), (3.0, 4.0)]:
        r = solve_hydrogen(N=48, L=L, Z=Z, n_steps=1500, order=4, verbose=False)
        assert abs((r.energy + 0.5 * Z * Z) / (0.5 * Z * Z)) < 0.08
        assert abs(r.virial_ratio - 1.0) < 0.08
