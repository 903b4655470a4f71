"""Orbital adjoints: the Rayleigh-quotient and RHF-energy gradients against
central finite differences of the same GPU functionals (FP64), and
Rayleigh-quotient descent against Lanczos."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_rayleigh_gradient_matches_finite_differences(gpu):
    import gato_amd as G
    grid = G.Grid3D(16, 8.0)
    H = G.Hamiltonian(grid, G.harmonic_oscillator(grid, 1.0), order=4)
    gen = torch.Generator(device="cuda").manual_seed(1)
    psi = torch.randn((16, 16, 16), generator=gen, dtype=torch.float64, device="cuda")
    d = torch.randn((16, 16, 16), generator=gen, dtype=torch.float64, device="cuda")
    x = psi.clone().requires_grad_(True)
    R = G.rayleigh_quotient(H, x)
    (g,) = torch.autograd.grad(R, x)
    eps = 1e-4
    fd = (float(G.rayleigh_quotient(H, psi + eps * d)) -
          float(G.rayleigh_quotient(H, psi - eps * d))) / (2 * eps)
    an = float((g * d).sum())
    assert an == pytest.approx(fd, rel=1e-7)
    # the gradient is orthogonal to psi (R is scale invariant)
    assert abs(float((g * psi).sum())) < 1e-12 * float(g.norm() * psi.norm())


@pytest.mark.parametrize("molecule", ["helium", "water"])
def test_rhf_energy_gradient_matches_finite_differences(gpu, molecule):
    """dE/dphi_i = 4 dV F phi_i: stencil, Hartree, exchange and (water) the
    HGH nonlocal term all enter the Fock apply of the backward pass."""
    import gato_amd as G
    N, L = 20, 10.0
    grid = G.Grid3D(N, L)
    if molecule == "helium":
        V, nl, n = G.softened_coulomb(grid, Z=2.0), None, 1
    else:
        from gato_amd.physics.water import water_hgh_elements, water_initial_geometry
        from gato_amd.pseudopotentials import nonlocal_plan
        nuc, el = water_initial_geometry(1.81, 104.5), water_hgh_elements()
        V, nl, n = G.hgh_local_potential(nuc, el, grid), nonlocal_plan(nuc, el, grid), 4
    gen = torch.Generator(device="cuda").manual_seed(2)
    stack = torch.randn((n, N, N, N), generator=gen, dtype=torch.float64, device="cuda") * 0.1
    d = torch.randn((n, N, N, N), generator=gen, dtype=torch.float64, device="cuda")
    x = stack.clone().requires_grad_(True)
    E = G.rhf_energy_ad(x, V, grid, V_nl_apply=nl)
    (g,) = torch.autograd.grad(E, x)
    def energy(s):
        return float(G.rhf_energy_ad(s.contiguous(), V, grid, V_nl_apply=nl))

    def central(eps):
        return (energy(stack + eps * d) - energy(stack - eps * d)) / (2 * eps)
    # the energy is quartic in the orbitals: Richardson-extrapolate the
    # central difference to remove its eps^2 term
    fd1, fd2 = central(2e-4), central(1e-4)
    fd = (4 * fd2 - fd1) / 3
    an = float((g * d).sum())
    print(molecule, an, fd1, fd2, fd, flush=True)
    assert an == pytest.approx(fd, rel=1e-8)


def test_rayleigh_descent_reaches_the_lanczos_ground_state(gpu):
    import gato_amd as G
    grid = G.Grid3D(24, 12.0)
    H = G.Hamiltonian(grid, G.harmonic_oscillator(grid, 1.0), order=2)
    gen = torch.Generator(device="cuda").manual_seed(4)
    psi0 = torch.randn((24, 24, 24), generator=gen, dtype=torch.float64, device="cuda")
    ref = G.lanczos(H, psi0, n_iters=80, n_eigenstates=1)
    psi, hist = G.rayleigh_descent(H, psi0, n_steps=800)
    energies = [e for _, e in hist]
    print(energies[:3], energies[-1], float(ref.eigenvalues[0]), flush=True)
    assert all(b <= a + 1e-12 for a, b in zip(energies, energies[1:]))
    assert energies[-1] == pytest.approx(float(ref.eigenvalues[0]), abs=1e-6)
    overlap = abs(float((psi * ref.eigenstates[..., 0]).sum()) * grid.dV)
    assert overlap == pytest.approx(1.0, abs=1e-5)
