"""The hot operator under HIP graph capture.

One Lanczos step of the H2O Fock operator (native Fock apply: stencil, HGH
nonlocal, exchange FFT solves; then the projection pass, the fused update
and the normalisation) is recorded into a graph on one stream and replayed
several times with new inputs, eager work and synchronisation between the
replays.  Every replay must equal the same step run eagerly, bit for bit:
the kernels record no memsets, the stencil takes its static variant rule
under capture (its variants agree bit for bit), and plans and workspaces
exist before capture from an eager warm-up on the capture stream."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,N", [(torch.float32, 48), (torch.float64, 32)])
def test_lanczos_step_graph_replays_match_eager(gpu, dtype, N):
    import gato_amd as G
    from gato_amd import _lib as L
    from gato_amd.physics.water import water_hgh_elements, water_initial_geometry
    from gato_amd.pseudopotentials import nonlocal_plan
    grid = G.Grid3D(N, 0.4 * N)
    nuc = water_initial_geometry(1.81, 104.5)
    el = water_hgh_elements()
    gen = torch.Generator(device="cuda").manual_seed(3)
    orb = torch.randn((N, N, N, 4), generator=gen, dtype=dtype, device="cuda")
    V = G.hgh_local_potential(nuc, el, grid)
    nl = nonlocal_plan(nuc, el, grid, torch.float64)
    F = G.RHFFock(grid, V, orb, order=4, V_nl_apply=nl)
    n3 = N ** 3
    code = L.dtype_code(dtype)
    nk = 6                                   # basis v_0 .. v_5, step j = 5
    lib = L.lib()
    rng = np.random.default_rng(5)

    def basis():
        # orthonormal in the grid inner product (CGS assumes it), built on the host
        q, _ = np.linalg.qr(rng.standard_normal((n3, nk)))
        q = q.T.reshape(nk, N, N, N) / np.sqrt(grid.dV)
        return torch.tensor(q, dtype=dtype, device="cuda")

    def step(Vb, w, coef, alpha, beta, nxt, ws, sp):
        w.copy_(F.apply(Vb[nk - 1]))
        L.check(lib.gato_cgs_coefficients(code, n3, nk, Vb.data_ptr(), n3, w.data_ptr(),
                                          grid.dV, coef.data_ptr(), alpha.data_ptr(), ws, sp))
        L.check(lib.gato_cgs_update(code, n3, nk, Vb.data_ptr(), n3, coef.data_ptr(),
                                    Vb[nk - 1].data_ptr(), alpha.data_ptr(), w.data_ptr(),
                                    w.data_ptr(), grid.dV, beta.data_ptr(), ws, sp))
        L.check(lib.gato_scale_by_inverse(code, n3, w.data_ptr(), beta.data_ptr(),
                                          nxt.data_ptr(), sp))

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        Vb = torch.empty((nk, N, N, N), dtype=dtype, device="cuda")
        w = torch.empty((N, N, N), dtype=dtype, device="cuda")
        nxt = torch.empty_like(w)
        coef = torch.empty(nk, dtype=torch.float64, device="cuda")
        alpha = torch.empty(1, dtype=torch.float64, device="cuda")
        beta = torch.empty(1, dtype=torch.float64, device="cuda")
        Vb.copy_(basis())
        ws = L.workspace().data_ptr()
        sp = L.stream_ptr()
        step(Vb, w, coef, alpha, beta, nxt, ws, sp)        # warm-up: plans, caches
    s.synchronize()

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        step(Vb, w, coef, alpha, beta, nxt, ws, sp)

    for rep in range(3):
        # new basis, eager work and a sync between replays
        Vb.copy_(basis())
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = (nxt.clone(), coef.clone(), alpha.clone(), beta.clone())
        with torch.cuda.stream(s):
            step(Vb, w, coef, alpha, beta, nxt, ws, sp)
        s.synchronize()
        want = (nxt, coef, alpha, beta)
        for a, b in zip(got, want):
            assert torch.equal(a, b), rep
        assert float(beta) > 0
        # the step's output is orthonormal to the basis (to working precision)
        d = (Vb.reshape(nk, -1).double() @ nxt.reshape(-1).double()) * grid.dV
        assert float(d.abs().max()) < (1e-5 if dtype == torch.float32 else 1e-12)
        assert abs(float((nxt.double() ** 2).sum()) * grid.dV - 1.0) < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "fp64", "mixed"])
def test_scf_with_graph_replayed_lanczos_equals_eager(gpu, precision):
    """The SCF loop's Lanczos runs replayed from the recorded graph (second
    iteration on) give the same energy history, orbital energies and orbitals
    as the eager native loop, bit for bit: H2O (HGH O/H) at 32^3, 6 SCF
    iterations from seeded random orbitals, with a full Fock apply and a
    device-wide synchronisation between the two runs."""
    import gato_amd as G
    from gato_amd.physics.water import water_hgh_elements, water_initial_geometry
    from gato_amd.pseudopotentials import nonlocal_plan
    from gato_amd.scf import scf_rhf
    N = 32
    grid = G.Grid3D(N, 0.4 * N)
    nuc = water_initial_geometry(1.81, 104.5)
    el = water_hgh_elements()
    V = G.hgh_local_potential(nuc, el, grid)
    nl = nonlocal_plan(nuc, el, grid, torch.float64)
    gen = torch.Generator(device="cuda").manual_seed(4)
    orb0 = torch.randn((N, N, N, 4), generator=gen, dtype=torch.float64, device="cuda")
    kw = dict(initial_orbitals=orb0, max_iters=6, tol=0.0, mixing=0.5, order=4, lanczos_iters=60,
              V_nl_apply=nl, precision=precision, switch_tol=1e3)
    eager = scf_rhf(V, grid, 4, use_graph=False, **kw)
    torch.cuda.synchronize()
    graphed = scf_rhf(V, grid, 4, use_graph=True, **kw)
    torch.cuda.synchronize()
    assert graphed.energy_history == eager.energy_history
    assert torch.equal(graphed.orbital_energies, eager.orbital_energies)
    assert torch.equal(graphed.orbitals, eager.orbitals)
    assert graphed.precision_history == eager.precision_history
