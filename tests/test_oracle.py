"""Pins the FP64 oracle before it is trusted as the checker.

Each oracle function is checked against what the reference's own tests pin
(closed forms, convergence rates, the He SCF constant; file:line under the
reference's tests/) or against what the mathematics fixes independently of
the oracle's formula: dense brute force on tiny grids (direct Coulomb sums,
dense numpy.linalg.eigh of the assembled Hamiltonian), and central finite
differences for the force.  These run on the CPU only.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

from oracle import gato_oracle as O


# --------------------------------------------------------------------------
# Laplacian (test_laplacian.py, test_laplacian_order4.py)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("order", [2, 4])
def test_periodic_plane_wave_is_exact_eigenvector(order):
    """test_laplacian.py:24-38 and test_laplacian_order4.py:26-57: cos(k.r)
    is an eigenvector of the periodic stencil with the closed-form symbol."""
    g = O.Grid3D(32, 2 * np.pi)
    X, Y, Z = g.coords()
    k = (1, 2, 3)
    psi = np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    h = g.h
    if order == 2:
        lam = sum((2 * np.cos(kk * h) - 2) / h ** 2 for kk in k)
    else:
        lam = sum((-2 * np.cos(2 * kk * h) + 32 * np.cos(kk * h) - 30) / (12 * h ** 2) for kk in k)
    assert np.max(np.abs(O.laplacian(psi, h, "periodic", order) - lam * psi)) < 1e-10


@pytest.mark.parametrize("order,rate", [(2, 2.0), (4, 4.0)])
def test_convergence_rate(order, rate):
    """test_laplacian.py:41-58, test_laplacian_order4.py:60-84: error of the
    symbol vs -|k|^2 falls as h^order."""
    errs = []
    for N in (16, 32):
        g = O.Grid3D(N, 2 * np.pi)
        X, Y, Z = g.coords()
        psi = np.sin(X + Y) * np.cos(Z)
        lap = O.laplacian(psi, g.h, "periodic", order)
        errs.append(np.max(np.abs(lap + 3 * psi)))
    assert abs(np.log2(errs[0] / errs[1]) - rate) < 0.1


@pytest.mark.parametrize("order", [2, 4])
def test_dirichlet_laplacian_is_hermitian(order):
    """test_laplacian.py:61-72: <a, L b> = <L a, b> (Dirichlet)."""
    g = O.Grid3D(12, 3.0)
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal((2, *g.shape))
    lhs = np.sum(a * O.laplacian(b, g.h, "dirichlet", order))
    rhs = np.sum(O.laplacian(a, g.h, "dirichlet", order) * b)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_dirichlet_matches_dense_matrix():
    """Brute force: the zero-padded stencil equals the Kronecker-sum matrix."""
    N, h = 5, 0.3
    D = (np.diag(np.full(N, -30.0)) + np.diag(np.full(N - 1, 16.0), 1) + np.diag(np.full(N - 1, 16.0), -1)
         + np.diag(np.full(N - 2, -1.0), 2) + np.diag(np.full(N - 2, -1.0), -2)) / (12 * h * h)
    I = np.eye(N)
    A = np.kron(np.kron(D, I), I) + np.kron(np.kron(I, D), I) + np.kron(np.kron(I, I), D)
    psi = np.random.default_rng(1).standard_normal((N, N, N))
    assert np.max(np.abs(O.laplacian(psi, h, "dirichlet", 4).ravel() - A @ psi.ravel())) < 1e-10


def test_gaussian_kinetic_energy():
    """test_laplacian_order4.py:101-109: <T> of exp(-r^2/2) is 3/4."""
    g = O.Grid3D(48, 12.0)
    X, Y, Z = g.coords()
    psi = O.normalize(np.exp(-(X * X + Y * Y + Z * Z) / 2), g)
    T = O.inner_product(psi, O.kinetic(psi, g.h, "dirichlet", 4), g)
    assert abs(T - 0.75) < 5e-3


def test_bad_order_and_boundary():
    with pytest.raises(ValueError):
        O.laplacian(np.zeros((4, 4, 4)), 0.1, "dirichlet", 3)
    with pytest.raises(ValueError):
        O.laplacian(np.zeros((4, 4, 4)), 0.1, "mirror", 2)


# --------------------------------------------------------------------------
# Hamiltonian, potentials, geometry
# --------------------------------------------------------------------------

def test_zero_potential_hamiltonian_is_kinetic():
    """test_hamiltonian.py:10-18."""
    g = O.Grid3D(16, 6.0)
    psi = np.random.default_rng(2).standard_normal(g.shape)
    H = O.Hamiltonian(g, np.zeros(g.shape))
    assert np.max(np.abs(H.apply(psi) - O.kinetic(psi, g.h))) < 1e-12


def test_softened_coulomb_and_multicenter():
    """test_potentials.py:8-46, test_multicenter_potential.py:11-73."""
    g = O.Grid3D(16, 8.0)
    V = O.softened_coulomb(g, Z=2.0, epsilon=0.1)
    X, Y, Z = g.coords()
    assert np.max(np.abs(V + 2.0 / np.sqrt(X * X + Y * Y + Z * Z + 0.01))) < 1e-14
    nuc1 = O.Nuclei(np.zeros((1, 3)), np.array([2.0]))
    assert np.max(np.abs(O.multi_center_softened_coulomb(nuc1, g, 0.1) - V)) < 1e-14
    nuc2 = O.diatomic_along_z(1.4, 1.0)
    V2 = O.multi_center_softened_coulomb(nuc2, g)
    Va = O.softened_coulomb(g, 1.0, center=(0, 0, -0.7))
    Vb = O.softened_coulomb(g, 1.0, center=(0, 0, 0.7))
    assert np.max(np.abs(V2 - Va - Vb)) < 1e-14


def test_nuclear_repulsion_and_force():
    """test_geometry.py:46-80: E_nn = 1/R, F = +-1/R^2."""
    nuc = O.diatomic_along_z(2.0, 1.0)
    assert abs(O.nuclear_repulsion(nuc) - 0.5) < 1e-15
    g = O.nuclear_repulsion_grad(nuc)
    assert abs(g[0, 2] - 0.25) < 1e-15 and abs(g[1, 2] + 0.25) < 1e-15


# --------------------------------------------------------------------------
# Poisson (test_poisson.py) and a brute-force convolution
# --------------------------------------------------------------------------

def _gauss(g, a, c=(0, 0, 0)):
    X, Y, Z = g.coords()
    return (a / np.pi) ** 1.5 * np.exp(-a * ((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2))


def test_gaussian_hartree_energy():
    """test_poisson.py:27-38: E_H = 1/2 sqrt(2a/pi) within 1 %."""
    g = O.Grid3D(48, 10.0)
    E = O.hartree_energy(_gauss(g, 0.5), g)
    ex = 0.5 * math.sqrt(2 * 0.5 / math.pi)
    assert abs(E - ex) / ex < 1e-2


def test_hartree_far_field():
    """test_poisson.py:41-57."""
    g = O.Grid3D(80, 10.0)
    J = O.hartree_potential(_gauss(g, 0.5), g)
    r = g.radial()
    mask = (r > 2.5) & (r < 3.5)
    assert np.max(np.abs(J[mask] - erf(np.sqrt(0.5) * r[mask]) / r[mask])) < 5e-4


def test_hartree_linear_positive_and_scaling():
    """test_poisson.py:60-90."""
    g = O.Grid3D(32, 8.0)
    r1, r2 = _gauss(g, 0.8, (0.5, 0, 0)), _gauss(g, 1.2, (-0.7, 0.3, 0))
    J = O.hartree_potential(1.3 * r1 - 0.4 * r2, g)
    assert np.max(np.abs(J - 1.3 * O.hartree_potential(r1, g) + 0.4 * O.hartree_potential(r2, g))) < 1e-10
    assert np.min(O.hartree_potential(_gauss(g, 1.0), g)) > 0
    g2 = O.Grid3D(48, 10.0)
    ratio = O.hartree_energy(_gauss(g2, 2.0), g2) / O.hartree_energy(_gauss(g2, 0.5), g2)
    assert abs(ratio - 2.0) / 2.0 < 2e-2


def test_hartree_equals_direct_coulomb_sum():
    """Brute force: J_i = sum_j rho_j / sqrt(|r_i - r_j|^2 + eps^2) dV."""
    g = O.Grid3D(5, 2.5)
    rho = np.random.default_rng(3).standard_normal(g.shape)
    X, Y, Z = (a.ravel() for a in g.coords())
    D2 = (X[:, None] - X[None]) ** 2 + (Y[:, None] - Y[None]) ** 2 + (Z[:, None] - Z[None]) ** 2
    eps = g.h / 2
    J = (1.0 / np.sqrt(D2 + eps * eps)) @ rho.ravel() * g.dV
    assert np.max(np.abs(O.hartree_potential(rho, g).ravel() - J)) < 1e-12 * np.max(np.abs(J))


# --------------------------------------------------------------------------
# Lanczos (test_lanczos.py) and dense eigh
# --------------------------------------------------------------------------

def test_lanczos_ho_ground_state_and_orthonormal():
    """test_lanczos.py:22-36, :70-92."""
    g = O.Grid3D(24, 10.0)
    H = O.Hamiltonian(g, O.harmonic_oscillator(g), order=4)
    X, Y, Z = g.coords()
    psi0 = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2) * (1 + 0.1 * (X + Y + Z))
    r = O.lanczos(H, psi0, 40, 3)
    assert abs(r.eigenvalues[0] - 1.5) < 1e-2
    S = np.moveaxis(r.eigenstates, -1, 0).reshape(3, -1)
    assert np.max(np.abs(S @ S.T * g.dV - np.eye(3))) < 1e-6


def test_lanczos_matches_dense_eigh():
    """Tiny grid: Lanczos with full reorthogonalisation over the whole space
    reproduces numpy.linalg.eigh of the assembled H."""
    g = O.Grid3D(5, 4.0)
    V = O.harmonic_oscillator(g)
    H = O.Hamiltonian(g, V, order=4)
    n = 125
    A = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        A[:, i] = H.apply(e.reshape(g.shape)).ravel()
    w = np.linalg.eigvalsh(A)
    psi0 = np.random.default_rng(4).standard_normal(g.shape)
    r = O.lanczos(H, psi0, 60, 3)
    assert np.max(np.abs(r.eigenvalues - w[:3])) < 1e-8


def test_default_krylov_dim_calibration():
    """test_lanczos.py:115-151."""
    dims = [O.default_krylov_dim(N, 4) for N in (32, 48, 64, 96, 128)]
    assert dims == sorted(dims) and dims[-1] > dims[0]
    assert O.default_krylov_dim(64, 4) >= 80 and O.default_krylov_dim(96, 4) >= 120
    assert O.default_krylov_dim(128, 4) == 160 and O.default_krylov_dim(160, 4) == 200
    assert O.default_krylov_dim(40, 4) >= 120 and O.default_krylov_dim(8, 12) > 12


# --------------------------------------------------------------------------
# Exchange, Fock, energy (test_scf.py)
# --------------------------------------------------------------------------

def _norm_gauss(g, a):
    X, Y, Z = g.coords()
    return O.normalize((2 * a / np.pi) ** 0.75 * np.exp(-a * (X * X + Y * Y + Z * Z)), g)


def test_exchange_reduces_to_self_for_one_orbital():
    """test_scf.py:44-53."""
    g = O.Grid3D(24, 8.0)
    phi = _norm_gauss(g, 1.0)
    K = O.exchange_apply(phi, phi[..., None], g)
    assert np.max(np.abs(K - phi * O.hartree_potential(phi * phi, g))) < 1e-12


def test_fock_with_zero_orbitals_is_core_hamiltonian():
    """test_scf.py:56-69."""
    g = O.Grid3D(16, 6.0)
    V = O.softened_coulomb(g, Z=2.0)
    fock = O.RHFFock(g, V, np.zeros((*g.shape, 1)), np.zeros(g.shape), order=4)
    psi = _norm_gauss(g, 1.5)
    assert np.max(np.abs(fock.apply(psi) - O.Hamiltonian(g, V, order=4).apply(psi))) < 1e-12


def test_single_orbital_energy_decomposition():
    """test_scf.py:72-94."""
    g = O.Grid3D(24, 8.0)
    phi = _norm_gauss(g, 1.0)
    E = O.rhf_energy(phi[..., None], np.zeros(g.shape), g)
    rho1 = phi * phi
    e2 = np.sum(rho1 * O.hartree_potential(rho1, g)) * g.dV
    e1 = 2 * O.inner_product(phi, O.kinetic(phi, g.h, "dirichlet", 4), g)
    assert abs(E - (e1 + e2)) < 1e-10


def test_exchange_matches_dense_two_electron_sum():
    """Brute force on a tiny grid: (K psi)_i = sum_j phi_j,i sum_k G_ik phi_j,k psi_k dV."""
    g = O.Grid3D(4, 2.0)
    rng = np.random.default_rng(5)
    orb = rng.standard_normal((*g.shape, 2))
    psi = rng.standard_normal(g.shape)
    X, Y, Z = (a.ravel() for a in g.coords())
    D2 = (X[:, None] - X[None]) ** 2 + (Y[:, None] - Y[None]) ** 2 + (Z[:, None] - Z[None]) ** 2
    G = 1.0 / np.sqrt(D2 + (g.h / 2) ** 2) * g.dV
    want = sum(orb[..., j].ravel() * (G @ (orb[..., j].ravel() * psi.ravel())) for j in range(2))
    assert np.max(np.abs(O.exchange_apply(psi, orb, g).ravel() - want)) < 1e-12


@pytest.mark.slow
def test_helium_scf_constant():
    """test_scf.py:97-126: He N=64 L=10 order 4 mixing 0.7 lanczos 40 gives
    E = -2.599 +- 0.010 and converges."""
    g = O.Grid3D(64, 10.0)
    r = O.scf_rhf(O.softened_coulomb(g, Z=2.0), g, 1, max_iters=25, tol=1e-6, mixing=0.7,
                  order=4, lanczos_iters=40)
    assert r.converged
    assert abs(r.energy + 2.599) < 0.010


# --------------------------------------------------------------------------
# HGH (test_pseudopotentials.py) and forces (test_water.py)
# --------------------------------------------------------------------------

def test_hgh_local_tail_and_origin():
    """test_pseudopotentials.py:39-113: tail -> -Z_ion/r; value at r=0 is
    -Z_ion sqrt(2/pi)/r_loc + C1."""
    p = O.lookup("O")
    g = O.Grid3D(16, 40.0)
    nuc = O.Nuclei(np.array([[0.3, 0.1, -0.2]]), np.array([8.0]))
    V = O.hgh_local_potential(nuc, (p,), g)
    gp = O._grid_pts(g)
    r = np.linalg.norm(gp - nuc.positions[0], axis=-1)
    far = r > 8
    assert np.max(np.abs(V[far] + p.Z_ion / r[far])) < 1e-6
    v0 = O._hgh_local_one_center(np.zeros((1, 3)), np.zeros(3), p)[0]
    assert abs(v0 - (-p.Z_ion * math.sqrt(2 / math.pi) / p.r_loc + p.c[0])) < 1e-9


@pytest.mark.parametrize("sym", ["O", "Li", "Cl"])
def test_projector_radial_normalised(sym):
    """test_pseudopotentials.py:120-138: int p(r)^2 r^2 dr = 1."""
    from scipy.integrate import quad
    for proj in O.lookup(sym).projectors:
        for i in range(1, proj.n + 1):
            val, _ = quad(lambda r: O.projector_radial(r, proj.ell, i, proj.r) ** 2 * r * r, 0, 50)
            assert abs(val - 1.0) < 1e-4


@pytest.mark.parametrize("sym", ["O", "Cl"])
def test_nonlocal_is_hermitian(sym):
    """test_pseudopotentials.py:145-167, :199-224."""
    g = O.Grid3D(16, 6.0)
    nuc = O.Nuclei(np.array([[0.1, -0.2, 0.05]]), np.array([8.0]))
    el = (O.lookup(sym),)
    rng = np.random.default_rng(6)
    a, b = rng.standard_normal((2, *g.shape))
    lhs = O.inner_product(a, O.hgh_nonlocal_apply(b, nuc, el, g), g)
    rhs = O.inner_product(O.hgh_nonlocal_apply(a, nuc, el, g), b, g)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def _hgh_projector_closed_form(x, y, z, ell, i_idx, r_l, axis=None):
    """The HGH 1998 projector p_i^l(r) Y_lm written out for a centre at the
    origin (Hartwigsen, Goedecker and Hutter, PRB 58, 3641, eq. 3), without
    the oracle's code: s = p(r)/sqrt(4 pi); p_x, p_y, p_z = p(r) sqrt(3/4pi)
    x/r etc.  ``axis`` picks the Cartesian factor for l = 1."""
    r2 = x * x + y * y + z * z
    n = ell + (4 * i_idx - 1) / 2.0
    pref = math.sqrt(2.0) / (r_l ** n * math.sqrt(math.gamma(n)))
    g = np.exp(-r2 / (2 * r_l * r_l)) * r2 ** (i_idx - 1)
    if ell == 0:
        return pref * g / math.sqrt(4 * math.pi)
    xyz = {"x": x, "y": y, "z": z}[axis]
    return pref * g * xyz * math.sqrt(3 / (4 * math.pi))


def _hgh_gram_closed_form(ell, i, j):
    """<p_i^l | p_j^l> over R^3 (angular part normalised to 1):
    int_0^inf p_i p_j r^2 dr = Gamma(l+i+j-1/2) / sqrt(Gamma(l+2i-1/2) Gamma(l+2j-1/2)),
    from int_0^inf r^(2a) e^(-r^2/r_l^2) dr = r_l^(2a+1) Gamma(a+1/2)/2."""
    return math.gamma(ell + i + j - 0.5) / math.sqrt(math.gamma(ell + 2 * i - 0.5)
                                                    * math.gamma(ell + 2 * j - 0.5))


@pytest.mark.parametrize("sym", ["O", "Li", "Cl"])
def test_nonlocal_self_projector_identity(sym):
    """V_nl applied to one of its own projector fields returns
    sum_i p_i (h G)_{i a}, with G the projectors' Gram matrix in closed form.
    The atom sits at the centre of a cell-centred grid, so projectors of
    different (l, m) are orthogonal on the grid by parity, and the grid is
    fine enough (h = r_l / 3 or finer) that the grid Gram matrix equals the
    continuous one to ~1e-12.  A missing dV, a scaled h_ij, a dropped
    off-diagonal h_12 (Cl s block) or an overlap taken without the radial
    Gram factor each fail this."""
    p = O.lookup(sym)
    for proj in p.projectors:
        r_l = proj.r
        L = 16.0 * r_l + 2.0 * r_l * (proj.ell + 2 * (proj.n - 1))
        N = int(math.ceil(L / (r_l / 3.0)))
        N += N % 2
        g = O.Grid3D(N, L)
        X, Y, Z = g.coords()
        nuc = O.Nuclei(np.zeros((1, 3)), np.array([1.0]))
        h = np.asarray(proj.h, dtype=float)
        G = np.array([[_hgh_gram_closed_form(proj.ell, i, j) for j in range(1, proj.n + 1)]
                      for i in range(1, proj.n + 1)])
        axes = [None] if proj.ell == 0 else ["y", "z", "x"]
        for ax in axes:
            fields = [_hgh_projector_closed_form(X, Y, Z, proj.ell, i, r_l, ax)
                      for i in range(1, proj.n + 1)]
            # the other channels of the atom must not respond: restrict the
            # element to this one projector set
            el = (O.HGHParams(p.symbol, p.Z_ion, p.r_loc, p.c, (proj,)),)
            for a in range(proj.n):
                out = O.hgh_nonlocal_apply(fields[a], nuc, el, g)
                want = sum(fields[i] * (h @ G)[i, a] for i in range(proj.n))
                scale = np.max(np.abs(want))
                assert scale > 0
                assert np.max(np.abs(out - want)) < 1e-9 * scale, (sym, proj.ell, ax, a)
        # the full element: the other channels of the same atom are orthogonal
        # to this projector set on the symmetric grid, so nothing else responds
        fields = [_hgh_projector_closed_form(X, Y, Z, proj.ell, i, r_l, axes[-1])
                  for i in range(1, proj.n + 1)]
        out = O.hgh_nonlocal_apply(fields[0], nuc, (p,), g)
        want = sum(fields[i] * (h @ G)[i, 0] for i in range(proj.n))
        assert np.max(np.abs(out - want)) < 1e-9 * np.max(np.abs(want)), (sym, proj.ell)


def test_nonlocal_l1_channel_order():
    """The l = 1 channels are m = -1 -> y, m = 0 -> z, m = +1 -> x
    (pseudopotentials.py:289-296).  A Li-centred field x f(r) has zero overlap
    with the m = -1 and m = 0 projectors and the closed-form overlap with the
    m = +1 one: for f = exp(-r^2/(2 r_l^2)), <p_x | x f> =
    N sqrt(3/4pi) int x^2 exp(-r^2/r_l^2) d^3r = N sqrt(3/4pi) pi^(3/2) r_l^5 / 2
    with N = sqrt(2) / (r_l^(5/2) sqrt(Gamma(5/2))).  V_nl of that field is
    h_11 times that overlap times the p_x projector."""
    p = O.lookup("Li")
    proj = [q for q in p.projectors if q.ell == 1][0]
    r_l = proj.r
    L = 18.0 * r_l
    g = O.Grid3D(64, L)
    X, Y, Z = g.coords()
    gp = O._grid_pts(g)
    f = X * np.exp(-(X * X + Y * Y + Z * Z) / (2 * r_l * r_l))
    Nrm = math.sqrt(2.0) / (r_l ** 2.5 * math.sqrt(math.gamma(2.5)))
    ov_x = Nrm * math.sqrt(3 / (4 * math.pi)) * math.pi ** 1.5 * r_l ** 5 / 2
    ovs = {m: O.inner_product(O.projector_field(gp, np.zeros(3), 1, m, 1, r_l), f, g)
           for m in (-1, 0, 1)}
    assert abs(ovs[-1]) < 1e-12 * ov_x and abs(ovs[0]) < 1e-12 * ov_x
    assert abs(ovs[1] - ov_x) < 1e-10 * ov_x
    nuc = O.Nuclei(np.zeros((1, 3)), np.array([3.0]))
    el = (O.HGHParams("Li", p.Z_ion, p.r_loc, p.c, (proj,)),)
    out = O.hgh_nonlocal_apply(f, nuc, el, g)
    want = proj.h[0][0] * ov_x * _hgh_projector_closed_form(X, Y, Z, 1, 1, r_l, "x")
    assert np.max(np.abs(out - want)) < 1e-9 * np.max(np.abs(want))


# Published HGH-LDA parameters (Hartwigsen, Goedecker, Hutter 1998, table I;
# the CP2K GTH-PADE q1/q6/q7 sets): Z_ion, r_loc, C1..C4, then per channel
# (l, r_l, h matrix).  pseudopotentials.py:122-174 tabulates the same numbers.
HGH_PUBLISHED = {
    "H": (1.0, 0.2, (-4.0663326, 0.6778322, 0.0, 0.0), ()),
    "O": (6.0, 0.24762086, (-16.5803180, 2.39570952, 0.0, 0.0),
          ((0, 0.22178614, ((18.26691718,),)),)),
    "Li": (1.0, 0.78755305, (-1.89261247, 0.28605968, 0.0, 0.0),
           ((0, 0.66637518, ((1.85881111,),)), (1, 1.07930561, ((-0.00589504,),)))),
    "Cl": (7.0, 0.41, (-6.86475431, 0.0, 0.0, 0.0),
           ((0, 0.33820832, ((9.06223968, -1.96193036), (-1.96193036, 5.06568240))),
            (1, 0.37613709, ((4.46587640,),)))),
}


@pytest.mark.parametrize("sym", sorted(HGH_PUBLISHED))
def test_hgh_tables_are_the_published_parameters(sym):
    """Both the oracle's and the product's tables hold the published values."""
    from gato_amd import pseudopotentials as P
    Zion, rloc, c, chans = HGH_PUBLISHED[sym]
    for p in (O.lookup(sym), P.lookup(sym)):
        assert (p.Z_ion, p.r_loc, tuple(p.c)) == (Zion, rloc, c)
        got = tuple((q.ell, q.r, tuple(tuple(r) for r in q.h)) for q in p.projectors)
        assert got == chans


def test_lookup_unknown_element():
    with pytest.raises(KeyError):
        O.lookup("Xx")


@pytest.mark.parametrize("elements", [("O", "H", "H"), ("Li", "H"), ("H", "Cl")])
def test_hf_force_matches_finite_differences(elements):
    """test_water.py:52-80 (delta = 1e-3): the analytic force of the oracle
    equals -dE/dR by central differences of its own BO energy, here on random
    orbitals (the identity holds for any fixed orbitals)."""
    g = O.Grid3D(16, 8.0)
    el = tuple(O.lookup(s) for s in elements)
    if len(el) == 3:
        nuc = O.water_initial_geometry(1.81, 104.5)
    else:
        nuc = O.diatomic_along_z(2.6, 1.0, 1.0)
    rng = np.random.default_rng(9)
    orb = rng.standard_normal((*g.shape, 2)) * 0.3
    F = O.hellmann_feynman_force(nuc.positions, nuc.charges, orb, el, g)
    d = 1e-4
    for k in range(len(el)):
        for c in range(3):
            Pp = nuc.positions.copy()
            Pm = nuc.positions.copy()
            Pp[k, c] += d
            Pm[k, c] -= d
            fd = -(O.bo_energy(Pp, nuc.charges, orb, el, g) - O.bo_energy(Pm, nuc.charges, orb, el, g)) / (2 * d)
            assert abs(fd - F[k, c]) < 1e-6 * max(1.0, abs(fd))


def test_adam_first_step_is_lr_sign():
    """Adam's bias correction makes the first step -lr * sign(g) (up to eps)."""
    a = O.Adam(0.05)
    u = a.update(np.array([[3.0, -2.0, 1e-3]]))
    assert np.allclose(u, -0.05 * np.sign([[3.0, -2.0, 1e-3]]), atol=1e-6)


# Kingma & Ba (2015), Algorithm 1, with b1 = 9/10, b2 = 999/1000, eps = 1e-8,
# lr = 1/10, worked by hand for the gradients g1 = 1, g2 = 3, g3 = -2:
#   t=1: m = 1/10, v = 1/1000; mhat = 1, vhat = 1            -> -lr/(1 + eps)
#   t=2: m = 39/100, v = 9999/10^6; mhat = 39/19,
#        vhat = 9999/1999                                   -> -lr (39/19)/(sqrt(9999/1999) + eps)
#   t=3: m = 151/1000, v = 13989001/10^9;
#        mhat = (151/1000)/(271/1000) = 151/271,
#        vhat = (13989001/10^9)/(1 - 0.999^3) = 13989001/2997001
#                                                           -> -lr (151/271)/(sqrt(13989001/2997001) + eps)
# With b1 and b2 swapped the first step is the same but the second is
# -lr (3999/1999)/(sqrt(99/19) + eps) = -0.0876..., not -0.0918...
ADAM_HAND = [-0.1 / (1 + 1e-8),
             -0.1 * (39 / 19) / (math.sqrt(9999 / 1999) + 1e-8),
             -0.1 * (151 / 271) / (math.sqrt(13989001 / 2997001) + 1e-8)]


def test_adam_three_steps_match_hand_derived_values():
    a = O.Adam(0.1)
    got = [float(a.update(np.array([g]))[0]) for g in (1.0, 3.0, -2.0)]
    assert np.allclose(got, ADAM_HAND, rtol=1e-13, atol=0)
    swapped = O.Adam(0.1, b1=0.999, b2=0.9)
    s = [float(swapped.update(np.array([g]))[0]) for g in (1.0, 3.0)]
    assert abs(s[1] - ADAM_HAND[1]) > 1e-3


def test_product_adam_matches_hand_derived_values():
    """The product's host-side Adam (positions are a (K, 3) host tensor)
    follows the same three hand-derived steps, elementwise per coordinate."""
    import torch
    from gato_amd.geometry_opt import Adam
    a = Adam(0.1)
    got = []
    for gv in (1.0, 3.0, -2.0):
        g = torch.full((2, 3), gv, dtype=torch.float64)
        g[1] *= -1.0  # negated gradients give negated updates
        u = a.update(g)
        assert torch.equal(u[1], -u[0])
        got.append(float(u[0, 0]))
    assert np.allclose(got, ADAM_HAND, rtol=1e-13, atol=0)


# --------------------------------------------------------------------------
# Imaginary time, LCAO guess, H2+ and hydrogen (solvers/imag_time.py,
# ansatz/grid.py, physics/h2_plus.py, physics/hydrogen.py)
# --------------------------------------------------------------------------

def test_imaginary_time_reaches_dense_ground_state():
    """Euler imaginary time projects onto the ground state of the assembled
    dense H (numpy.linalg.eigh): energy and state (up to sign).  A sign error in
    the step would climb to the top of the spectrum instead."""
    g = O.Grid3D(6, 6.0)
    H = O.Hamiltonian(g, O.harmonic_oscillator(g), order=2)
    n = g.N ** 3
    A = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        A[:, i] = H.apply(e.reshape(g.shape)).ravel()
    w, U = np.linalg.eigh(A)
    psi0 = np.ones(g.shape) + 0.1 * np.random.default_rng(5).standard_normal(g.shape)
    psi, hist = O.imaginary_time(H, psi0, n_steps=400, log_every=100)
    assert abs(hist.energies[-1] - w[0]) < 1e-10
    assert hist.steps == [0, 100, 200, 300, 399]
    assert all(a >= b - 1e-12 for a, b in zip(hist.energies, hist.energies[1:]))
    u = U[:, 0] / np.sqrt(g.dV)
    s = np.sign(np.dot(u, psi.ravel()))
    assert np.max(np.abs(s * u - psi.ravel())) < 1e-6


def test_init_lcao_is_a_normalisable_1s_per_centre():
    """Z^{3/2} exp(-Z r) integrates to pi for any Z (a wrong prefactor or decay
    constant changes it by powers of Z)."""
    g = O.Grid3D(64, 8.0)
    nuc = O.Nuclei(np.zeros((1, 3)), np.array([2.0]))
    p = O.init_lcao(nuc, g, softening=0.02)
    assert abs(O.norm_sq(p, g) - math.pi) < 0.03 * math.pi
    two = O.init_lcao(O.diatomic_along_z(2.0, 1.0), O.Grid3D(16, 8.0))
    assert np.allclose(two, two[:, :, ::-1])  # homonuclear: symmetric under z -> -z


def test_h2p_force_matches_finite_differences():
    """Analytic Hellmann-Feynman force of H2+ at fixed psi against central
    differences of the BO energy (test_h2_plus.py:54-76 pattern, delta 1e-4)."""
    g = O.Grid3D(16, 8.0)
    nuc = O.diatomic_along_z(2.2, 1.0)
    pt = O.h2p_solve_electronic(nuc, g, n_steps=200)
    eps = g.h / 2
    F = O.h2p_hellmann_feynman_force(nuc.positions, nuc.charges, pt.psi, g, eps)
    d = 1e-4
    for k, a in [(1, 2), (0, 2), (1, 0)]:
        Pp, Pm = nuc.positions.copy(), nuc.positions.copy()
        Pp[k, a] += d
        Pm[k, a] -= d
        fd = -(O.h2p_bo_energy(Pp, nuc.charges, pt.psi, g, eps) -
               O.h2p_bo_energy(Pm, nuc.charges, pt.psi, g, eps)) / (2 * d)
        assert abs(fd - F[k, a]) < 1e-6 * max(1.0, abs(fd))
    assert abs(F[0, 2] + F[1, 2]) < 1e-8  # symmetric pair: forces cancel
    assert abs(pt.energy - (pt.electronic_energy + pt.nuclear_repulsion)) < 1e-14


def test_hydrogen_imag_time_energy_and_virial():
    """Coarse-grid hydrogen (test_hydrogen.py:13-21 bounds, scaled down):
    softening-limited energy just above -1/2 and virial ratio near 1."""
    r = O.solve_hydrogen(N=24, L=10.0, n_steps=600, order=4)
    assert -0.5 < r.energy < -0.40
    assert abs(r.virial_ratio - 1.0) < 0.1
    assert abs(r.energy - (r.kinetic + r.potential)) < 1e-14
