"""GPU parity of the Lanczos, SCF, energy and force paths against the FP64
oracle at sizes the oracle finishes in seconds."""
import numpy as np
import pytest
import torch

from oracle import gato_oracle as O

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def host(t):
    return t.detach().double().cpu().numpy()


def test_lanczos_harmonic_oscillator_matches_oracle(gpu):
    import gato_amd as G
    g, gg = O.Grid3D(24, 10.0), G.Grid3D(24, 10.0)
    rng = np.random.default_rng(7)
    psi0 = rng.standard_normal(g.shape)
    ref = O.lanczos(O.Hamiltonian(g, O.harmonic_oscillator(g), order=4), psi0, 60, 6)
    H = G.Hamiltonian(gg, G.harmonic_oscillator(gg), order=4)
    res = G.lanczos(H, dev(psi0), n_iters=60, n_eigenstates=6)
    assert np.max(np.abs(res.eigenvalues.numpy() - ref.eigenvalues)) < 1e-9
    # alpha, beta ~ 35: the one-sweep classical vs modified Gram-Schmidt
    # difference is rounding-level per step and grows over 60 steps (~1e-7)
    assert np.max(np.abs(res.alphas - ref.alphas)) < 1e-6
    assert np.max(np.abs(res.betas - ref.betas)) < 1e-6
    # Ritz vectors orthonormal (test_lanczos.py:70-92) and equal to the oracle's up to sign
    S = host(res.states).reshape(6, -1)
    ov = S @ S.T * g.dV
    assert np.max(np.abs(ov - np.eye(6))) < 1e-8
    R = np.moveaxis(ref.eigenstates, -1, 0).reshape(6, -1)
    for k in range(3):
        s = np.sign(np.dot(S[k], R[k]))
        assert np.max(np.abs(s * S[k] - R[k])) < 1e-6


def test_lanczos_fp32_basis(gpu):
    import gato_amd as G
    gg = G.Grid3D(24, 10.0)
    rng = np.random.default_rng(8)
    psi0 = rng.standard_normal(gg.shape)
    H = G.Hamiltonian(gg, G.harmonic_oscillator(gg), order=4)
    r64 = G.lanczos(H, dev(psi0), 60, 4)
    r32 = G.lanczos(H, dev(psi0, torch.float32), 60, 4)
    assert abs(float(r32.eigenvalues[0]) - float(r64.eigenvalues[0])) < 1e-5


def test_generic_operator_lanczos_matches_native(gpu):
    import gato_amd as G
    gg = G.Grid3D(20, 10.0)
    H = G.Hamiltonian(gg, G.harmonic_oscillator(gg), order=2)

    class Wrapped:
        grid = gg

        def apply(self, psi):
            return H.apply(psi)

    psi0 = dev(np.random.default_rng(3).standard_normal(gg.shape))
    a = G.lanczos(H, psi0, 40, 3)
    b = G.lanczos(Wrapped(), psi0, 40, 3)
    assert np.max(np.abs(a.eigenvalues.numpy() - b.eigenvalues.numpy())) < 1e-10


def test_helium_scf_matches_oracle(gpu):
    import gato_amd as G
    N, L = 24, 8.0
    g, gg = O.Grid3D(N, L), G.Grid3D(N, L)
    ref = O.scf_rhf(O.softened_coulomb(g, Z=2.0), g, 1, max_iters=12, tol=1e-9, mixing=0.7,
                    order=4, lanczos_iters=40)
    res = G.scf_rhf(G.softened_coulomb(gg, Z=2.0), gg, 1, max_iters=12, tol=1e-9, mixing=0.7,
                    order=4, lanczos_iters=40)
    assert res.n_iters == ref.n_iters
    assert abs(res.energy - ref.energy) < 1e-9
    assert np.max(np.abs(np.array(res.energy_history) - np.array(ref.energy_history))) < 1e-8


def _water_ref(N, L, iters):
    g = O.Grid3D(N, L)
    nuc = O.water_initial_geometry(1.81, 104.5)
    el = O.water_hgh_elements()
    res, E, Enn = O.solve_rhf_at_geometry(nuc, el, g, max_iters=iters, tol=1e-10, mixing=0.5,
                                          order=4)
    return g, nuc, el, res, E


@pytest.fixture(scope="module")
def water_small():
    return _water_ref(32, 12.0, 6)


def test_water_scf_fp64_matches_oracle(gpu, water_small):
    import gato_amd as G
    from gato_amd.physics.water import solve_rhf_at_geometry, water_hgh_elements
    g, nuc_o, el_o, ref, Eref = water_small
    gg = G.Grid3D(32, 12.0)
    nuc = G.make_nuclei(nuc_o.positions, nuc_o.charges)
    p = solve_rhf_at_geometry(nuc, water_hgh_elements(), gg, max_iters=6, tol=1e-10, mixing=0.5)
    hist = np.array(p.energy_history)
    assert np.max(np.abs(hist - np.array(ref.energy_history))) < 1e-8
    assert np.max(np.abs(p.orbital_energies.numpy() - ref.orbital_energies)) < 1e-8


def test_water_scf_mixed_precision_reaches_fp64(gpu):
    import gato_amd as G
    from gato_amd.physics.water import solve_rhf_at_geometry, water_hgh_elements
    gg = G.Grid3D(32, 12.0)
    nuc_o = O.water_initial_geometry(1.81, 104.5)
    nuc = G.make_nuclei(nuc_o.positions, nuc_o.charges)
    p64 = solve_rhf_at_geometry(nuc, water_hgh_elements(), gg, max_iters=40, tol=1e-8, mixing=0.5)
    pmx = solve_rhf_at_geometry(nuc, water_hgh_elements(), gg, max_iters=40, tol=1e-8, mixing=0.5,
                                precision="mixed")
    assert p64.converged and pmx.converged
    assert abs(pmx.energy - p64.energy) < 1e-6


def test_water_force_matches_oracle_and_fd(gpu, water_small):
    import gato_amd as G
    from gato_amd.physics.water import water_hgh_elements
    g, nuc_o, el_o, ref, Eref = water_small
    gg = G.Grid3D(32, 12.0)
    orb = ref.orbitals
    F_ref = O.hellmann_feynman_force(nuc_o.positions, nuc_o.charges, orb, el_o, g)
    F = G.hellmann_feynman_force(torch.tensor(nuc_o.positions), torch.tensor(nuc_o.charges),
                                 dev(orb), water_hgh_elements(), gg)
    assert np.max(np.abs(F.numpy() - F_ref)) < 1e-8
    # central finite difference of the GPU BO energy, delta = 1e-3 (test_water.py:52-80)
    d = 1e-3
    for (k, c) in [(0, 2), (1, 0), (2, 2)]:
        P = torch.tensor(nuc_o.positions)
        Pp, Pm = P.clone(), P.clone()
        Pp[k, c] += d
        Pm[k, c] -= d
        Ep = float(G.bo_energy(Pp, torch.tensor(nuc_o.charges), dev(orb), water_hgh_elements(), gg))
        Em = float(G.bo_energy(Pm, torch.tensor(nuc_o.charges), dev(orb), water_hgh_elements(), gg))
        fd = -(Ep - Em) / (2 * d)
        assert abs(fd - float(F[k, c])) < 1e-3 * abs(fd) + 1e-4
