"""Reductions over more Krylov vectors than one pass of the reduction
workspace holds partials for (16 MiB of double partials: 151 FP64 vectors of
192^3 in the projection pass).  The basis is then processed in slabs; these
tests run such sizes through the C ABI and the native Lanczos loop."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N = 192


@pytest.fixture(scope="module")
def basis(gpu):
    g = torch.Generator(device="cuda").manual_seed(11)
    V = torch.randn((320, N ** 3), generator=g, device="cuda", dtype=torch.float64)
    w = torch.randn(N ** 3, generator=g, device="cuda", dtype=torch.float64)
    yield V, w
    del V, w
    torch.cuda.empty_cache()


def test_projection_coefficients_in_slabs_match_double_reference(basis):
    """gato_cgs_coefficients with 240 FP64 vectors of 192^3 (two slabs):
    alpha = s<v_j, w>, coef_k = s<v_k, w> - alpha s<v_k, v_j>."""
    from gato_amd import _lib as L
    V, w = basis
    nk, n, scale = 240, N ** 3, 0.37
    print(f"projection: nk={nk} n={n} FP64", flush=True)
    coef = torch.full((nk,), float("nan"), device="cuda", dtype=torch.float64)
    alpha = torch.full((1,), float("nan"), device="cuda", dtype=torch.float64)
    L.check(L.lib().gato_cgs_coefficients(L.dtype_code(torch.float64), n, nk, V.data_ptr(), n,
                                          w.data_ptr(), scale, coef.data_ptr(), alpha.data_ptr(),
                                          L.workspace(V.device).data_ptr(), L.stream_ptr()))
    D = V[:nk] @ w
    G = V[:nk] @ V[nk - 1]
    a = scale * D[nk - 1]
    want = scale * D - a * (scale * G)
    size = scale * D.abs() + a.abs() * scale * G.abs()
    torch.cuda.synchronize()
    assert abs(float(alpha[0] - a)) < 1e-11 * float(scale * D[nk - 1].abs() + 1)
    err = ((coef - want).abs() / size).max()
    print("max relative error", float(err), flush=True)
    assert float(err) < 1e-11


def test_broadcast_dots_in_slabs_match_double_reference(basis):
    """gato_dot_batched of 320 FP64 vectors of 192^3 against one vector: the
    partials of all 320 do not fit, so the vectors go in slabs."""
    from gato_amd import _native as nat
    V, w = basis
    print("broadcast dots: nk=320 FP64", flush=True)
    got = nat.dot_batched(V, w, 320, N ** 3, 0, scale=0.5)
    want = 0.5 * (V @ w)
    torch.cuda.synchronize()
    assert float(((got - want).abs() / (0.5 * (V.abs() @ w.abs()))).max()) < 1e-12


def _separable_well(n, L, depth=2.0, width=1.0):
    h = L / n
    x = -L / 2 + (np.arange(n) + 0.5) * h
    return x, h, -depth * np.exp(-x * x / (2 * width * width))


def test_fp64_lanczos_192_past_the_single_pass_limit(gpu):
    """240 FP64 Lanczos steps on 192^3 (every projection past step 150 runs in
    slabs).  The potential v(x) + v(y) + v(z) is separable, so the order-2
    Dirichlet Hamiltonian is a Kronecker sum and its lowest eigenvalue is
    three times the lowest eigenvalue of the 192 x 192 one-dimensional
    matrix, from numpy.linalg.eigvalsh."""
    import gato_amd as G
    L = 48.0
    x, h, v = _separable_well(N, L)
    T1 = (np.diag(np.full(N, 1.0 / h ** 2)) + np.diag(np.full(N - 1, -0.5 / h ** 2), 1)
          + np.diag(np.full(N - 1, -0.5 / h ** 2), -1) + np.diag(v))
    lam = 3 * np.linalg.eigvalsh(T1)[0]
    vt = torch.tensor(v, dtype=torch.float64, device="cuda")
    Vf = (vt[:, None, None] + vt[None, :, None] + vt[None, None, :]).contiguous()
    gg = G.Grid3D(N, L)
    H = G.Hamiltonian(gg, Vf, order=2)
    g = torch.Generator(device="cuda").manual_seed(5)
    psi0 = torch.randn((N, N, N), generator=g, device="cuda", dtype=torch.float64)
    print(f"Lanczos FP64 {N}^3 m=240, exact {lam:.12f}", flush=True)
    res = G.lanczos(H, psi0, n_iters=240, n_eigenstates=2)
    e0 = float(res.eigenvalues[0])
    print("Lanczos", e0, "error", e0 - lam, flush=True)
    assert abs(e0 - lam) < 1e-9
    S = res.states.reshape(2, -1)
    ov = (S @ S.T * gg.dV).cpu().numpy()
    assert np.max(np.abs(ov - np.eye(2))) < 1e-8
    del res, S
    torch.cuda.empty_cache()
