"""GPU parity of each HIP kernel against the FP64 oracle on the same seeded
inputs.  Tolerances: FP64 paths to ~1e-11 relative (rounding order differs
from numpy/pocketfft), FP32 paths to ~1e-5 relative."""
import numpy as np
import pytest
import torch

from oracle import gato_oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(1234)


def dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def host(t):
    return t.detach().double().cpu().numpy()


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# N % 4 == 0 (float) / N % 2 == 0 (double) take the 16-byte vector kernel:
# 2 and 4 wrap inside one vector, 136 spans two z-tiles and a partial y-tile;
# 33 (and 70 in float) take the LDS-tiled scalar kernel
@pytest.mark.parametrize("N", [2, 4, 8, 24, 33, 70, 136])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("boundary", ["dirichlet", "periodic"])
def test_laplacian_matches_oracle(gpu, N, order, boundary):
    import gato_amd as G
    psi = RNG.standard_normal((N, N, N))
    h = 0.37
    ref = O.laplacian(psi, h, boundary, order)
    got = host(G.laplacian(dev(psi), h, boundary, order))
    assert relerr(got, ref) < 1e-13
    got32 = host(G.laplacian(dev(psi, torch.float32), h, boundary, order))
    assert relerr(got32, ref) < 5e-6


@pytest.mark.parametrize("N", [20, 21])
def test_hamiltonian_batched_and_fused_terms(gpu, N):
    import gato_amd as G
    from gato_amd import _native as nat
    h = 0.3
    psi = RNG.standard_normal((3, N, N, N))
    V = RNG.standard_normal((N, N, N))
    prev = RNG.standard_normal((3, N, N, N))
    base = RNG.standard_normal((3, N, N, N))
    coef = torch.tensor([0.7], dtype=torch.float64, device="cuda")
    out = dev(base)
    nat.stencil(dev(psi), h, 4, "dirichlet", -0.5, V=dev(V), out=out, accumulate=True,
                sub=dev(prev), sub_coef=coef)
    ref = np.stack([O.kinetic(psi[b], h, "dirichlet", 4) + V * psi[b] + base[b] - 0.7 * prev[b]
                    for b in range(3)])
    assert relerr(host(out), ref) < 1e-13
    out32 = dev(base, torch.float32)
    nat.stencil(dev(psi, torch.float32), h, 4, "dirichlet", -0.5, V=dev(V, torch.float32),
                out=out32, accumulate=True, sub=dev(prev, torch.float32), sub_coef=coef)
    assert relerr(host(out32), ref) < 5e-6


# large batches: 64^3 float x 64 fields, 70^3 double x 20 (70 % 4 != 0: the
# last row pair of a y-tile is partial), fused V, accumulate and subtract
# both batches take the x-marching kernel (>= 512 blocks of 32-plane chunks);
# 70^3 double has partial tiles in y and z and a 6-plane last chunk
@pytest.mark.parametrize("N,batch,dtype", [(64, 64, torch.float32), (70, 20, torch.float64)])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("boundary", ["dirichlet", "periodic"])
def test_stencil_large_batch(gpu, N, batch, dtype, order, boundary):
    from gato_amd import _native as nat
    h = 0.41
    rng = np.random.default_rng(N + batch + order)
    psi = rng.standard_normal((batch, N, N, N))
    V = rng.standard_normal((N, N, N))
    base = rng.standard_normal((batch, N, N, N))
    prev = rng.standard_normal((batch, N, N, N))
    coef = torch.tensor([0.3], dtype=torch.float64, device="cuda")
    out = dev(base, dtype)
    nat.stencil(dev(psi, dtype), h, order, boundary, -0.5, V=dev(V, dtype), out=out,
                accumulate=True, sub=dev(prev, dtype), sub_coef=coef)
    got = host(out)
    tol = 1e-13 if dtype == torch.float64 else 5e-6
    for k in (0, batch // 2, batch - 1):
        ref = (-0.5 * O.laplacian(psi[k], h, boundary, order) + V * psi[k] + base[k]
               - 0.3 * prev[k])
        assert relerr(got[k], ref) < tol, k


def test_bad_order_and_boundary_raise(gpu):
    import gato_amd as G
    psi = dev(np.zeros((4, 4, 4)))
    with pytest.raises(ValueError):
        G.laplacian(psi, 0.1, "dirichlet", 3)
    with pytest.raises(ValueError):
        G.laplacian(psi, 0.1, "mirror", 2)


@pytest.mark.parametrize("n", [1, 1000, 262144 + 17])
def test_dot_and_norm(gpu, n):
    from gato_amd import _native as nat
    a = RNG.standard_normal(n)
    b = RNG.standard_normal(n)
    got = float(nat.dot_batched(dev(a), dev(b), 1, 0, 0, scale=0.5)[0])
    ref = 0.5 * float(np.dot(a, b))
    assert abs(got - ref) <= 1e-12 * max(1.0, np.sum(np.abs(a * b)))


def test_batched_dot_and_cgs(gpu):
    import ctypes as C
    from gato_amd import _lib as L, _native as nat
    n, k = 5000, 37
    V = RNG.standard_normal((k, n))
    w = RNG.standard_normal(n)
    got = host(nat.dot_batched(dev(V), dev(w), k, n, 0, scale=1.0))
    assert np.max(np.abs(got - V @ w)) < 1e-10
    S0 = RNG.standard_normal(n)
    s0 = torch.tensor([1.7], dtype=torch.float64, device="cuda")
    o = torch.empty(k, dtype=torch.float64, device="cuda")
    Vd_, wd_, Sd_ = dev(V), dev(w), dev(S0)  # keep the operands alive across the launch
    L.check(L.lib().gato_dot_batched(L.F64, n, k, Vd_.data_ptr(), n, wd_.data_ptr(), 0, None,
                                     Sd_.data_ptr(), s0.data_ptr(), 1.0, o.data_ptr(),
                                     L.workspace().data_ptr(), L.stream_ptr()))
    assert np.max(np.abs(host(o) - V @ (w - 1.7 * S0))) < 1e-10
    c = RNG.standard_normal(k)
    wd, Vd, cd = dev(w), dev(V), dev(c)
    beta = torch.empty(1, dtype=torch.float64, device="cuda")
    out = torch.empty_like(wd)
    S = RNG.standard_normal(n)
    sc = torch.tensor([0.3], dtype=torch.float64, device="cuda")
    Sd = dev(S)
    L.check(L.lib().gato_cgs_update(L.F64, n, k, Vd.data_ptr(), n, cd.data_ptr(), Sd.data_ptr(),
                                    sc.data_ptr(), wd.data_ptr(), out.data_ptr(), 2.0,
                                    beta.data_ptr(), L.workspace().data_ptr(), L.stream_ptr()))
    ref = w - 0.3 * S - c @ V
    assert relerr(host(out), ref) < 1e-13
    assert abs(float(beta[0]) - np.sqrt(2.0 * ref @ ref)) < 1e-10


@pytest.mark.parametrize("n", [5000, 5001, 2 * 128 ** 3])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_cgs_coefficients(gpu, n, dtype):
    """alpha = s<v_j, w>, c_k = s<v_k, w - alpha v_j> (v_j the last of k vectors) in one
    pass; n = 5001 takes the scalar path, 2*128^3 the chunked fast path."""
    from gato_amd import _lib as L
    k = 7
    s = 0.25
    V = RNG.standard_normal((k, n))
    V /= np.sqrt(s * np.sum(V * V, axis=1))[:, None]  # s <v_k, v_k> = 1, as in Lanczos
    w = RNG.standard_normal(n) / np.sqrt(s * n) + 3.0 * V[-1]
    Vd, wd = dev(V, dtype), dev(w, dtype)
    Vh, wh = host(Vd), host(wd)  # the values the kernel sees
    coef = torch.empty(k, dtype=torch.float64, device="cuda")
    alpha = torch.empty(1, dtype=torch.float64, device="cuda")
    L.check(L.lib().gato_cgs_coefficients(L.dtype_code(dtype), n, k, Vd.data_ptr(), n,
                                          wd.data_ptr(), s, coef.data_ptr(), alpha.data_ptr(),
                                          L.workspace().data_ptr(), L.stream_ptr()))
    a_ref = s * float(Vh[-1] @ wh)
    c_ref = s * (Vh @ (wh - a_ref * Vh[-1]))
    # double accumulation: error bounded by rounding of the summed magnitudes
    bound = 1e-13 * s * (np.abs(Vh) @ (np.abs(wh) + abs(a_ref) * np.abs(Vh[-1])))
    assert abs(float(alpha[0]) - a_ref) < 1e-13 * s * float(np.abs(Vh[-1]) @ np.abs(wh))
    assert np.all(np.abs(host(coef) - c_ref) < bound)


@pytest.mark.parametrize("n", [1, 4097, 2 * 128 ** 3])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_normalize_matches_oracle(gpu, n, dtype):
    """gato_normalize: v = w / sqrt(s<w, w>) and the norm, out of place and in
    place; n = 4097 takes the scalar path, 2*128^3 the vector path with 2048
    wave partials summed by every block; a zero vector stays zero."""
    from gato_amd import _lib as L
    from gato_amd import _native as nat
    s = 0.125
    w = RNG.standard_normal(n) * 3.0
    wd = dev(w, dtype)
    wh = host(wd)
    ref_norm = np.sqrt(s * float(wh @ wh))
    norm = torch.empty(1, dtype=torch.float64, device="cuda")
    v = nat.normalize(wd, s, norm=norm)
    tol = 1e-14 if dtype == torch.float64 else 1e-7
    assert abs(float(norm[0]) - ref_norm) < 1e-14 * ref_norm
    assert np.max(np.abs(host(v) - wh / ref_norm)) < tol * np.max(np.abs(wh / ref_norm))
    # in place, twice in a row: the second call sees a unit vector
    nat.normalize(wd, s, out=wd, norm=norm)
    assert np.array_equal(host(wd), host(v))
    nat.normalize(wd, s, out=wd, norm=norm)
    assert abs(float(norm[0]) - 1.0) < (1e-14 if dtype == torch.float64 else 1e-6)
    # NULL norm pointer goes to the workspace slot; zero input gives zeros
    z = torch.zeros(n, dtype=dtype, device="cuda")
    out = torch.full_like(z, 7.0)
    L.check(L.lib().gato_normalize(L.dtype_code(dtype), n, z.data_ptr(), s, None, out.data_ptr(),
                                   L.workspace().data_ptr(), L.stream_ptr()))
    assert not torch.any(out != 0).item()


# 32 and 128 take the four-step kernels (2N = 8^2, 16^2); the others the
# generic mixed-radix passes (radix 3, 5, 7 included)
@pytest.mark.parametrize("N", [6, 16, 20, 24, 28, 32, 128])
def test_hartree_potential_matches_oracle(gpu, N):
    from gato_amd.solvers.poisson import PoissonPlan
    g = O.Grid3D(N, 0.4 * N)
    rho = RNG.standard_normal((3, N, N, N))
    ref = np.stack([O.hartree_potential(rho[k], g) for k in range(3)])
    plan = PoissonPlan(N, g.h, g.h / 2, torch.float64)
    got = host(plan.hartree_potential(dev(rho)))
    assert relerr(got, ref) < 1e-11
    E = host(plan.hartree_energy(dev(rho)))
    Eref = np.array([0.5 * O.integrate(rho[k] * ref[k], g) for k in range(3)])
    assert np.max(np.abs(E - Eref)) < 1e-10 * np.max(np.abs(Eref))
    plan32 = PoissonPlan(N, g.h, g.h / 2, torch.float32)
    got32 = host(plan32.hartree_potential(dev(rho, torch.float32)))
    assert relerr(got32, ref) < 2e-5


@pytest.mark.parametrize("N", [20, 32, 128])
def test_exchange_and_pair_integrals_match_oracle(gpu, N):
    from gato_amd.solvers.poisson import PoissonPlan
    g = O.Grid3D(N, 0.4 * N)
    orb = RNG.standard_normal((N, N, N, 3))
    psi = RNG.standard_normal((N, N, N))
    ref = O.exchange_apply(psi, orb, g)
    plan = PoissonPlan(N, g.h, g.h / 2, torch.float64)
    stack = dev(np.moveaxis(orb, -1, 0))
    got = host(plan.exchange_apply(dev(psi), stack, 1.0))
    assert relerr(got, ref) < 1e-11
    plan32 = PoissonPlan(N, g.h, g.h / 2, torch.float32)
    got32 = host(plan32.exchange_apply(dev(psi, torch.float32),
                                       dev(np.moveaxis(orb, -1, 0), torch.float32), 1.0))
    assert relerr(got32, ref) < 2e-5
    # a field that is not 16-byte aligned takes the generic z passes
    buf = torch.zeros(N ** 3 + 1, dtype=torch.float32, device="cuda")
    psi_mis = buf[1:].view(N, N, N)
    psi_mis.copy_(dev(psi, torch.float32))
    assert psi_mis.data_ptr() % 16 != 0
    got_mis = host(plan32.exchange_apply(psi_mis, dev(np.moveaxis(orb, -1, 0), torch.float32), 1.0))
    assert relerr(got_mis, ref) < 2e-5
    pairs = host(plan.pair_integrals(stack))
    k = 0
    for i in range(3):
        for j in range(i, 3):
            dens = orb[..., i] * orb[..., j]
            want = O.integrate(dens * O.hartree_potential(dens, g), g)
            assert abs(pairs[k] - want) < 1e-10 * abs(want) + 1e-12
            k += 1


# 2N = 128, 160, 192 and 320 take the four-step kernels of shape 16 x 8, 16 x 10,
# 16 x 12 and 16 x 20 (and, for a field that is not 16-byte aligned, the
# compiled in-place Stockham z passes); five orbitals give three exchange
# pairs per launch, the last one half used
@pytest.mark.parametrize("N,n_orb", [(64, 5), (80, 5), (96, 5), (160, 3)])
def test_compiled_fft_sizes_match_oracle(gpu, N, n_orb):
    from gato_amd.solvers.poisson import PoissonPlan
    g = O.Grid3D(N, 0.2 * N)
    rho = RNG.standard_normal((N, N, N))
    ref = O.hartree_potential(rho, g)
    plan = PoissonPlan(N, g.h, g.h / 2, torch.float64)
    plan32 = PoissonPlan(N, g.h, g.h / 2, torch.float32)
    assert relerr(host(plan.hartree_potential(dev(rho))), ref) < 1e-11
    assert relerr(host(plan32.hartree_potential(dev(rho, torch.float32))), ref) < 2e-5
    # the energy epilogue (block partials of int rho J) of the same shapes
    Eref = 0.5 * O.integrate(rho * ref, g)
    assert abs(float(host(plan.hartree_energy(dev(rho[None])))[0]) - Eref) < 1e-10 * abs(Eref)
    E32 = float(host(plan32.hartree_energy(dev(rho[None], torch.float32)))[0])
    assert abs(E32 - Eref) < 2e-5 * abs(Eref)
    orb = RNG.standard_normal((N, N, N, n_orb))
    psi = RNG.standard_normal((N, N, N))
    ref = O.exchange_apply(psi, orb, g)
    stack = np.moveaxis(orb, -1, 0)
    got = host(plan.exchange_apply(dev(psi), dev(stack), 1.0))
    assert relerr(got, ref) < 1e-11
    got32 = host(plan32.exchange_apply(dev(psi, torch.float32), dev(stack, torch.float32), 1.0))
    assert relerr(got32, ref) < 2e-5
    buf = torch.zeros(N ** 3 + 1, dtype=torch.float32, device="cuda")
    psi_mis = buf[1:].view(N, N, N)
    psi_mis.copy_(dev(psi, torch.float32))
    assert psi_mis.data_ptr() % 16 != 0
    got_mis = host(plan32.exchange_apply(psi_mis, dev(stack, torch.float32), 1.0))
    assert relerr(got_mis, ref) < 2e-5


# every stencil variant (one-plane kernel, marching chunks of 4-64 planes with
# 1-3 planes of prefetch, in both block orders, with LDS tiles or barrier-free
# wave strips) and the autotuned and static choices give the same bits
@pytest.mark.parametrize("N,batch,dtype,boundary", [(64, 64, torch.float32, "dirichlet"),
                                                    (128, 1, torch.float32, "dirichlet"),
                                                    (96, 6, torch.float32, "dirichlet"),
                                                    (96, 4, torch.float32, "periodic"),
                                                    (70, 20, torch.float64, "periodic")])
def test_stencil_variants_bitwise_identical(gpu, N, batch, dtype, boundary):
    import gato_amd as G
    from gato_amd import _lib as L
    from gato_amd import _native as nat
    rng = np.random.default_rng(N + batch)
    psi = dev(rng.standard_normal((batch, N, N, N)), dtype)
    V = dev(rng.standard_normal((N, N, N)), dtype)
    acc0 = dev(rng.standard_normal((batch, N, N, N)), dtype)
    sub = dev(rng.standard_normal((batch, N, N, N)), dtype)
    coef = torch.tensor([0.37], dtype=torch.float64, device="cuda")

    def run():
        out = acc0.clone()
        nat.stencil(psi, 0.3, 4, boundary, -0.5, V=V, out=out, accumulate=True, sub=sub,
                    sub_coef=coef)
        torch.cuda.synchronize()
        return out
    lib = L.lib()
    try:
        ref = run()                       # autotuned
        G.set_autotune(False)
        assert torch.equal(run(), ref)    # static rule
        for v in (0, 4, 8, 16, 32, 104, 132, 208, 216, 232, 264, 1008, 1016, 1032, 1132, 1216,
                  10032, 11008, 11116, 100008, 100232, 101016, 110032, 111008, 111116, 111208,
                  200008, 201016, 200116, 211032, 221032, 221164):
            lib.gato_stencil_force_variant(v)
            assert torch.equal(run(), ref), v
    finally:
        lib.gato_stencil_force_variant(-1)
        G.set_autotune(True)
    want = O.kinetic(host(psi[0]), 0.3, boundary, 4) + host(V) * host(psi[0]) + host(acc0[0]) \
        - 0.37 * host(sub[0])
    assert relerr(host(ref[0]), want) < (1e-12 if dtype == torch.float64 else 2e-5)


# the same for the plain Laplacian (no potential, accumulate or subtract: the
# buffer-load variants take those inputs under uniform branches)
@pytest.mark.parametrize("N,batch,dtype", [(64, 16, torch.float32), (48, 8, torch.float64)])
def test_stencil_variants_bitwise_identical_plain(gpu, N, batch, dtype):
    import gato_amd as G
    from gato_amd import _lib as L
    from gato_amd import _native as nat
    rng = np.random.default_rng(3 * N + batch)
    psi = dev(rng.standard_normal((batch, N, N, N)), dtype)

    def run():
        out = nat.stencil(psi, 0.25, 4, "dirichlet", 1.0)
        torch.cuda.synchronize()
        return out
    lib = L.lib()
    try:
        ref = run()
        for v in (0, 8, 32, 1008, 11008, 100008, 100032, 101008, 111008, 111116, 111208,
                  201008, 211116, 220032):
            lib.gato_stencil_force_variant(v)
            assert torch.equal(run(), ref), v
    finally:
        lib.gato_stencil_force_variant(-1)
    want = O.laplacian(host(psi[1]), 0.25, "dirichlet", 4)
    assert relerr(host(ref[1]), want) < (1e-12 if dtype == torch.float64 else 2e-5)


# the fused Euler step + renormalisation of imaginary time against the stencil
# followed by gato_normalize; N = 15 does not take the vector path (fallback)
@pytest.mark.parametrize("N,dtype,boundary", [(32, torch.float64, "dirichlet"),
                                              (128, torch.float32, "dirichlet"),
                                              (24, torch.float32, "periodic"),
                                              (15, torch.float64, "dirichlet")])
def test_stencil_normalize_matches_two_step(gpu, N, dtype, boundary):
    from gato_amd import _lib as L
    from gato_amd import _native as nat
    rng = np.random.default_rng(N)
    psi = dev(rng.standard_normal((N, N, N)), dtype)
    V = dev(rng.standard_normal((N, N, N)), dtype)
    coef = torch.tensor([-1.0], dtype=torch.float64, device="cuda")
    dV, h, c = 0.027, 0.3, 0.004
    ref = nat.stencil(psi, h, 4, boundary, c, V=V, sub=psi, sub_coef=coef)
    nrm_ref = torch.empty(1, dtype=torch.float64, device="cuda")
    ref = nat.normalize(ref, dV, norm=nrm_ref)
    out = torch.empty_like(psi)
    nrm = torch.empty(1, dtype=torch.float64, device="cuda")
    L.check(L.lib().gato_stencil_normalize(L.dtype_code(dtype), N, h, 4, nat.boundary_code(boundary),
                                           c, psi.data_ptr(), V.data_ptr(), psi.data_ptr(),
                                           coef.data_ptr(), out.data_ptr(), dV, nrm.data_ptr(),
                                           L.workspace().data_ptr(), L.stream_ptr()))
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    assert relerr(host(out), host(ref)) < tol
    assert abs(float(nrm) - float(nrm_ref)) < 1e-12 * float(nrm_ref)


# the multi-step imaginary-time entry (renormalisation deferred into the next
# step's launch) against one fused step + renormalisation per step; N = 15
# takes the per-step fallback
@pytest.mark.parametrize("N,dtype,boundary,steps", [(32, torch.float64, "dirichlet", 7),
                                                    (32, torch.float64, "periodic", 6),
                                                    (128, torch.float32, "dirichlet", 9),
                                                    (24, torch.float32, "periodic", 1),
                                                    (15, torch.float64, "dirichlet", 5)])
def test_imag_time_steps_match_per_step(gpu, N, dtype, boundary, steps):
    from gato_amd import _lib as L
    from gato_amd import _native as nat
    rng = np.random.default_rng(100 + N)
    h, dV, dt = 0.3, 0.027, 0.009
    psi = nat.normalize(dev(1.0 + 0.3 * rng.standard_normal((N, N, N)), dtype), dV)
    Vdt = dev(-dt * rng.uniform(-2.0, 0.5, (N, N, N)), dtype)
    coef = torch.tensor([-1.0], dtype=torch.float64, device="cuda")
    lib, ws, sp, code = L.lib(), L.workspace().data_ptr(), L.stream_ptr(), L.dtype_code(dtype)
    bnd = nat.boundary_code(boundary)
    ref = psi.clone()
    nrm_ref = torch.empty(1, dtype=torch.float64, device="cuda")
    tmp = torch.empty_like(psi)
    for _ in range(steps):
        L.check(lib.gato_stencil_normalize(code, N, h, 4, bnd, 0.5 * dt, ref.data_ptr(),
                                           Vdt.data_ptr(), ref.data_ptr(), coef.data_ptr(),
                                           tmp.data_ptr(), dV, nrm_ref.data_ptr(), ws, sp))
        ref, tmp = tmp, ref
    a, b = psi.clone(), torch.empty_like(psi)
    nrm = torch.empty(1, dtype=torch.float64, device="cuda")
    L.check(lib.gato_imag_time_steps(code, N, h, 4, bnd, dt, Vdt.data_ptr(), a.data_ptr(),
                                     b.data_ptr(), steps, dV, nrm.data_ptr(), ws, sp))
    got = b if steps % 2 else a
    tol = 1e-13 if dtype == torch.float64 else 1e-5
    assert relerr(host(got), host(ref)) < tol
    assert abs(float(nrm) - float(nrm_ref)) < tol * float(nrm_ref)
    assert abs(float((got.double() ** 2).sum()) * dV - 1.0) < (1e-12 if dtype == torch.float64
                                                                else 1e-5)


def test_local_potentials_match_oracle(gpu):
    import gato_amd as G
    g = O.Grid3D(24, 9.0)
    gg = G.Grid3D(24, 9.0)
    ref = O.softened_coulomb(g, Z=2.0, center=(0.3, -0.2, 0.1))
    got = host(G.softened_coulomb(gg, Z=2.0, center=(0.3, -0.2, 0.1)))
    assert relerr(got, ref) < 1e-14
    nuc_o = O.water_initial_geometry(1.81, 104.5)
    nuc = G.make_nuclei(nuc_o.positions, nuc_o.charges)
    ref = O.hgh_local_potential(nuc_o, O.water_hgh_elements(), g)
    from gato_amd.physics.water import water_hgh_elements
    got = host(G.hgh_local_potential(nuc, water_hgh_elements(), gg))
    assert relerr(got, ref) < 1e-13
    ref = O.harmonic_oscillator(g, 1.3)
    got = host(G.harmonic_oscillator(gg, 1.3))
    assert relerr(got, ref) < 1e-14


# boxes of at most 8192 points take the single-pass kernel, larger ones the
# overlap pass plus the reducing add: water on the fine grid (O box of ~22000
# points), LiH (the Li p box covers the whole grid; at 96^3 the overlap pass
# is capped at 512 blocks) and HCl at 96^3 (Cl box of ~85000 points)
@pytest.mark.parametrize("molecule,N,L", [("water", 28, 12.0), ("hcl", 28, 12.0),
                                          ("lih", 28, 12.0), ("water", 64, 8.0),
                                          ("lih", 96, 14.0), ("hcl", 96, 14.0)])
def test_nonlocal_apply_matches_oracle(gpu, molecule, N, L):
    import gato_amd as G
    from gato_amd.physics import molecules as M
    from gato_amd.physics.water import water_hgh_elements
    g = O.Grid3D(N, L)
    gg = G.Grid3D(N, L)
    if molecule == "water":
        nuc_o, el_o, el = O.water_initial_geometry(1.81, 104.5), O.water_hgh_elements(), water_hgh_elements()
    elif molecule == "hcl":
        nuc_o = O.diatomic_along_z(2.41, 1.0, 17.0)
        el_o, el = (O.lookup("H"), O.lookup("Cl")), M.hcl_hgh_elements()
    else:
        nuc_o = O.diatomic_along_z(3.02, 3.0, 1.0)
        el_o, el = (O.lookup("Li"), O.lookup("H")), M.lih_hgh_elements()
    nuc = G.make_nuclei(nuc_o.positions, nuc_o.charges)
    psi = RNG.standard_normal((2, N, N, N))
    ref = np.stack([O.hgh_nonlocal_apply(psi[k], nuc_o, el_o, g) for k in range(2)])
    plan = G.NonlocalPP(nuc, el, gg, torch.float64)
    if N == 64:  # only O carries projectors: its box alone
        assert plan.n_points > 8192
    got = host(plan(dev(psi)))
    assert relerr(got, ref) < 1e-11
    E = host(plan.expectation(dev(psi)))
    Eref = np.array([O.inner_product(psi[k], ref[k], g) for k in range(2)])
    assert np.max(np.abs(E - Eref)) < 1e-10 * np.max(np.abs(Eref))
    if N >= 64:  # the float plan of the FP32 iterations
        p32 = G.NonlocalPP(nuc, el, gg, torch.float32)
        got32 = host(p32(dev(psi, torch.float32)))
        assert relerr(got32, ref) < 1e-5


# Ritz combination out_k = sum_j S[j, k] V_j: the vectorised K <= 4 kernel and
# the general one (K = 6, or a length that is not a multiple of the vector
# width) against numpy in double
@pytest.mark.parametrize("dtype,n,m,K", [(torch.float32, 4096 * 3, 37, 4),
                                         (torch.float64, 4096 * 3, 20, 3),
                                         (torch.float32, 1000, 16, 1),
                                         (torch.float32, 1001, 9, 4),
                                         (torch.float64, 2048, 12, 6)])
def test_ritz_combine_matches_numpy(gpu, dtype, n, m, K):
    from gato_amd import _lib as L
    rng = np.random.default_rng(n + m + K)
    Vh = rng.standard_normal((m, n))
    Sh = rng.standard_normal((m, K))
    V = dev(Vh, dtype)
    S = torch.tensor(Sh, dtype=torch.float64, device="cuda").contiguous()
    out = torch.empty((K, n), dtype=dtype, device="cuda")
    L.check(L.lib().gato_ritz_combine(L.dtype_code(dtype), n, m, K, V.data_ptr(), n,
                                      S.data_ptr(), out.data_ptr(), n, L.stream_ptr()))
    want = Sh.T @ host(V)
    tol = 1e-14 if dtype == torch.float64 else 1e-6
    assert relerr(host(out), want) < tol


@pytest.mark.parametrize("N", [128, 160, 96, 80, 64])
def test_fft_pencil_choices_bitwise_identical(gpu, N):
    """The four-step passes with 32 and with 16 pencils per block (the two
    choices the plan times) give the same bits: exchange of 4 orbitals on
    float grids of every four-step shape (2N = 256, 320, 192, 160, 128)."""
    import gato_amd as G
    from gato_amd import _lib as L
    from gato_amd.solvers.poisson import get_plan
    grid = G.Grid3D(N, 0.2 * N)
    rng = np.random.default_rng(12)
    psi = dev(rng.standard_normal((N, N, N)), torch.float32)
    orb = dev(rng.standard_normal((4, N, N, N)), torch.float32)
    plan = get_plan(grid, None, torch.float32)
    lib = L.lib()
    outs = []
    try:
        for p in (32, 16):
            lib.gato_poisson_force_pencils(p)
            outs.append(plan.exchange_apply(psi, orb, 1.0).clone())
            torch.cuda.synchronize()
    finally:
        lib.gato_poisson_force_pencils(0)
    assert torch.equal(outs[0], outs[1])
