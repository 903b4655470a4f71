#!/usr/bin/env python
"""Headline benchmark: H2O RHF SCF iterations per second on one MI355X.

Workload (the H2O row of the benchmark definition): 128^3 grid, h = 0.2 bohr
(L = 25.6), HGH O (Z_ion 6, local + s projector) and two HGH H at O-H 1.81
bohr / 104.5 deg, 4 doubly occupied orbitals, 13-point (order 4) stencil,
softening eps = h/2 = 0.1, Krylov dimension 160 (= default_krylov_dim(128, 4)),
linear density mixing 0.5.  Orbitals start from seeded random fields.

A step is one SCF solve to convergence (|dE| < 1e-7 Ha by default) from the
seeded start.  ``value`` is the FP32-mixed path's SCF iterations per second
over whole solves.  Its float iterations keep fields, stencils, FFTs and the
Krylov basis in float (inner products, Lanczos coefficients and eigensolves
in double) and solve each iteration's eigenproblem inexactly: block Lanczos
seeded with the current orbitals, stopped once the Ritz residuals fall below
a tolerance that tightens with |dE|, on the Fock operator with the
compressed exchange of the current orbitals (exact on them, built from
their 16 exchange solves once per iteration).  When the float energies
converge, FP64 iterations of the same scheme, with Ritz residuals below
0.05 sqrt(tol), run until two consecutive ones differ by less than the
tolerance (same fixed point: the compressed exchange equals the exact one
on the orbitals).  The plain FP64 GPU path (the reference's algorithm in double:
scalar Lanczos over the whole 160-vector Krylov dimension, exact exchange)
is solved beside it from the same start, and so are the FP32-mixed solve
with FP64 iterations of the reference's algorithm, and the same precision
pattern with the reference's algorithm in every iteration (the gain of
precision alone); the line states the times to solution, and the total
energy, orbital energies,
virial ratio and forces of each solve with their errors against an FP64 solve
to |dE| < 1e-9.  All inputs are resident in HBM before timing starts.

The default run also carries the H2O 160^3 geometry optimisation (100 Adam
steps on Hellmann-Feynman forces, finite-difference force checks every 10
steps); ``--geomopt-only`` runs it alone and prints its own JSON line.

Also reported: the same solve with block Lanczos over the whole Krylov
dimension, the uneven batch
(independent H2, LiH and HCl molecules on 64^3 to 128^3 grids, each solved to
convergence on concurrent streams, aggregate SCF iterations per second), the
hydrogen 64^3 Lanczos and helium 96^3 FP64 SCF workloads, the fused-stencil
HBM bandwidth on a batch of 16 orbitals (and over 4-32 orbitals beside the
element-wise ceiling), the dominant kernel's roofline (live HIP-event timing of
its float launches in two further solves of the same workload, outside the
timed region), and the FP64
CPU oracle timed on this host on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "H2O RHF 128^3 SCF iters/s, FP32-mixed vs FP64 GPU @1e-6 Ha; stencil HBM GB/s"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_system(N: int, h: float):
    import torch
    import gato_amd as G
    from gato_amd.physics.water import water_hgh_elements, water_initial_geometry
    from gato_amd.pseudopotentials import nonlocal_plan
    grid = G.Grid3D(N, h * N)
    nuc = water_initial_geometry(1.81, 104.5)
    el = water_hgh_elements()
    V = G.hgh_local_potential(nuc, el, grid)
    nl = nonlocal_plan(nuc, el, grid, torch.float64)
    gen = torch.Generator(device="cuda").manual_seed(0)
    orb0 = torch.randn((4, N, N, N), generator=gen, device="cuda", dtype=torch.float64)
    return grid, nuc, el, V, nl, orb0


PROBE_KINDS = {
    1: "cgs_update_kernel (reorthogonalisation update w -= alpha v_j + V c; reads j+1 basis "
       "vectors + w + v_j, writes w)",
    2: "proj2_kernel (Lanczos projections: alpha_j = <v_j, w> and c = V^T (w - alpha_j v_j) in "
       "one pass; reads the j+1 basis vectors and w)",
    3: "fs_pass_kernel / fft_pass_kernel (strided y-forward, x with kernel multiply, y-inverse "
       "FFT passes of the Poisson solves)",
    4: "fs_zfwd_kernel + fs_zinv_kernel (z FFT passes with fused pair product / exchange "
       "epilogue)",
    5: "block_dot_staged_kernel (block-Lanczos projection D = V^T [Q_i, W] on the matrix cores, "
       "v_mfma_f64_4x4x4f64 on float fields widened to double; reads the nk basis vectors and "
       "the 2b columns)",
    6: "block_update_kernel (block-Lanczos update W -= V C; reads nk basis vectors and the b "
       "columns, writes the b columns)",
}

# committed rocprofv3 --pmc summaries per probe kind (HBM bytes per launch)
PMC_SUMMARIES = {2: "proj2_kernel_h2o128_pmc.md", 5: "block_dot_staged_h2o128_pmc.md"}


def solve_h2o(grid, V, nl, orb0, precision: str, tol: float, eigensolver: str = "auto"):
    """One SCF solve of the H2O workload from the seeded start to |dE| < tol;
    returns (SCFResult, wall seconds).  The wall time covers everything the
    call does: the start orbitals' conversion, every iteration (Hartree solve,
    Lanczos, mixing, energy), the FP32 -> FP64 switch of the mixed path."""
    import torch
    from gato_amd.scf import scf_rhf, to_grid
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = scf_rhf(V, grid, 4, initial_orbitals=to_grid(orb0), max_iters=200, tol=tol, mixing=0.5,
                order=4, epsilon=grid.h / 2, V_nl_apply=nl, precision=precision,
                eigensolver=eigensolver)
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def split_iterations(r) -> dict:
    """Iteration counts and steady per-iteration times of one solve, by precision."""
    ph = [p for p in r.precision_history if p != "fp64-eval"]
    t32 = [s for s, p in zip(r.iteration_seconds, ph) if p == "fp32"]
    t64 = [s for s, p in zip(r.iteration_seconds, ph) if p == "fp64"]

    def steady_ms(ts):
        ts = ts[2:] if len(ts) > 4 else ts   # the first iterations allocate and tune
        return None if not ts else round(1e3 * sorted(ts)[len(ts) // 2], 3)
    return {"fp32_iterations": len(t32), "fp64_iterations": len(t64),
            "fp32_iteration_ms": steady_ms(t32), "fp64_iteration_ms": steady_ms(t64)}


PROBE_SOLVES = 2


def time_solves(grid, V, nl, orb0, precision, tol, warmup: int, steps: int, probe: int = 0,
                eigensolver: str = "auto", label: str = ""):
    """``warmup`` untimed solves, then ``steps`` timed ones (one step = one
    solve to convergence).  The live probe, when asked, brackets every launch
    of its kernel kind on float fields with HIP events, which adds a few
    microseconds around each; so it runs over PROBE_SOLVES further solves of
    the same workload after the timed ones, outside the timed region."""
    from gato_amd import _lib as L
    import ctypes as C
    for w in range(warmup):
        r, dt = solve_h2o(grid, V, nl, orb0, precision, tol, eigensolver)
        log(f"  {label} warmup {w}: {r.n_iters} iterations in {dt:.3f} s, E = {r.energy:+.10f}")
    runs = []
    for s in range(steps):
        r, dt = solve_h2o(grid, V, nl, orb0, precision, tol, eigensolver)
        runs.append((r, dt))
        log(f"  {label} step {s}: {r.n_iters} iterations ({split_iterations(r)}) in {dt:.3f} s, "
            f"converged {r.converged}, E = {r.energy:+.10f}")
    stats = None
    if probe:
        L.lib().gato_probe_reset(int(probe) | (4 << 8))
        for _ in range(PROBE_SOLVES):
            r, dt = solve_h2o(grid, V, nl, orb0, precision, tol, eigensolver)
        cnt, ms, byt = C.c_long(), C.c_double(), C.c_double()
        L.check(L.lib().gato_probe_read(C.byref(cnt), C.byref(ms), C.byref(byt)))
        L.lib().gato_probe_reset(0)
        stats = (cnt.value, ms.value, byt.value)
        log(f"  {label}: {PROBE_SOLVES} probed solves for the roofline ({r.n_iters} iterations, "
            f"{dt:.3f} s each with the probe's events)")
    return runs, stats


def h2o_observables(r, grid, V, nl, nuc, el) -> dict:
    """The four observables of a converged H2O solve, all evaluated in FP64
    from its orbitals: total energy (with the valence-charge E_nn), the
    orbital energies (the last Lanczos run's Ritz values, as scf_rhf returns
    them, and the Rayleigh quotients <phi_i|F|phi_i> of the Fock operator of
    the final orbitals), the virial ratio 2<T>/|E - <T>| and the
    Hellmann-Feynman forces."""
    import torch
    from gato_amd import _native as nat
    from gato_amd.forces import hellmann_feynman_force
    from gato_amd.physics.water import pp_nuclear_repulsion
    from gato_amd.scf import RHFFock, decompose_rhf_energy, to_stack
    orb = r.orbitals if r.orbitals.dtype == torch.float64 else r.orbitals.double()
    terms = decompose_rhf_energy(orb, V, grid, order=4, epsilon=grid.h / 2, V_nl_apply=nl)
    E_nn = float(pp_nuclear_repulsion(nuc, el))
    E_tot = terms.total + E_nn
    F = RHFFock(grid=grid, V_ext=V, orbitals=orb, epsilon=grid.h / 2, V_nl_apply=nl, order=4)
    stack = to_stack(orb)
    rq = []
    for i in range(stack.shape[0]):
        Fp = F.apply(stack[i].contiguous())
        rq.append(float(nat.dot_batched(stack[i].contiguous(), Fp, 1, 0, 0, scale=grid.dV)[0]))
    forces = hellmann_feynman_force(nuc.positions, nuc.charges, orb, el, grid, 4)
    return {"energy_total": E_tot, "energy_rhf": float(r.energy),
            "orbital_energies_ritz": [float(x) for x in r.orbital_energies],
            "orbital_energies_rayleigh": rq,
            "virial_ratio": 2.0 * terms.kinetic / abs(E_tot - terms.kinetic),
            "forces": forces.detach().double().cpu().numpy().tolist()}


def compare_observables(a: dict, b: dict) -> dict:
    """Absolute differences of the observables of solve a against solve b."""
    import numpy as np
    return {"energy_total": abs(a["energy_total"] - b["energy_total"]),
            "orbital_energies_ritz_max": float(np.max(np.abs(
                np.subtract(a["orbital_energies_ritz"], b["orbital_energies_ritz"])))),
            "orbital_energies_rayleigh_max": float(np.max(np.abs(
                np.subtract(a["orbital_energies_rayleigh"], b["orbital_energies_rayleigh"])))),
            "virial_ratio": abs(a["virial_ratio"] - b["virial_ratio"]),
            "forces_max": float(np.max(np.abs(np.subtract(a["forces"], b["forces"]))))}


def stencil_bandwidth(N: int, h: float, reps: int = 20, variant: int = -1, batch: int = 16):
    """Fused T + V psi on a batch of float orbitals (16 by default): GB/s of
    algorithmic traffic (read the batch and V, write the batch).  variant >= 0
    forces that stencil variant (a profiled run then measures the same kernel as
    an unprofiled one); the default takes the autotuned choice."""
    import torch
    from gato_amd import _lib as L
    from gato_amd import _native as nat
    psi = torch.randn((batch, N, N, N), device="cuda", dtype=torch.float32)
    V = torch.randn((N, N, N), device="cuda", dtype=torch.float32)
    out = torch.empty_like(psi)
    if variant >= 0:
        L.lib().gato_stencil_force_variant(variant)
    try:
        for _ in range(3):
            nat.stencil(psi, h, 4, "dirichlet", -0.5, V=V, out=out)
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            nat.stencil(psi, h, 4, "dirichlet", -0.5, V=V, out=out)
        e.record()
        torch.cuda.synchronize()
    finally:
        if variant >= 0:
            L.lib().gato_stencil_force_variant(-1)
    ms = s.elapsed_time(e) / reps
    nbytes = (2 * batch + 1) * N ** 3 * 4
    return nbytes / (ms * 1e-3) / 1e9, ms


def pointwise_bandwidth(N: int, batch: int, reps: int = 20):
    """torch's element-wise out = V * psi over the same batch (V broadcast): the
    same algorithmic traffic as the fused stencil with no neighbours, the rate a
    one-pass kernel of that volume reaches on this part."""
    import torch
    psi = torch.randn((batch, N, N, N), device="cuda", dtype=torch.float32)
    V = torch.randn((N, N, N), device="cuda", dtype=torch.float32)
    out = torch.empty_like(psi)
    for _ in range(3):
        torch.mul(psi, V, out=out)
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        torch.mul(psi, V, out=out)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    return (2 * batch + 1) * N ** 3 * 4 / (ms * 1e-3) / 1e9


def stencil_batches(N: int, h: float, batches=(4, 8, 16, 32)):
    """The fused stencil (autotuned) and the element-wise ceiling over batch sizes:
    up to 8 fields of 128^3 (<= 143 MB) the repeated launches run from the 256 MB
    Infinity Cache, from 16 fields on they stream from HBM."""
    rows = []
    for b in batches:
        g, ms = stencil_bandwidth(N, h, batch=b)
        pw = pointwise_bandwidth(N, b)
        rows.append({"fields": b, "MB": round((2 * b + 1) * N ** 3 * 4 / 1e6, 1),
                     "stencil_GBps": round(g, 1), "stencil_frac_of_peak": round(g / HBM_PEAK_GBPS, 4),
                     "stencil_us": round(ms * 1e3, 2), "pointwise_GBps": round(pw, 1),
                     "stencil_over_pointwise": round(g / pw, 3)})
        log(f"  stencil {b:2d} fields: {g:6.0f} GB/s ({g / HBM_PEAK_GBPS:.3f} of peak), "
            f"element-wise V*psi {pw:6.0f} GB/s")
    return rows


def imag_time_rate(N: int, h: float, steps: int = 200, dtype_name: str = "float32"):
    """Imaginary-time steps/s on H2+ (R = 2 bohr) at N^3: one stencil launch
    per step (reads psi and V, writes out scaled by the inverse norm of psi
    and the block partials of <out, out>), i.e. 3 field streams per step,
    plus one normalisation pass at the end of the run."""
    import torch
    import gato_amd as G
    dtype = getattr(torch, dtype_name)
    grid = G.Grid3D(N, h * N)
    nuc = G.diatomic_along_z(2.0, 1.0)
    H = G.Hamiltonian(grid, G.multi_center_softened_coulomb(nuc, grid, dtype=dtype), order=4)
    psi0 = G.init_lcao(nuc, grid, dtype=dtype)
    G.imaginary_time(H, psi0, n_steps=5, log_every=10 ** 9)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    G.imaginary_time(H, psi0, n_steps=steps, log_every=10 ** 9)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s = torch.finfo(dtype).bits // 8
    return steps / dt, (3 * steps + 2) * s * N ** 3 / dt / 1e9


def scf_iteration_bytes(N: int, m: int, n_occ: int, s: int, nl_box: int = 18 ** 3) -> float:
    """Algorithmic HBM bytes of one SCF iteration of this implementation (each
    kernel's inputs read once and outputs written once; s = bytes per real).

    Per Lanczos step j (m steps): the Fock matvec = exchange (n_occ solves as
    n_occ/2 complex pairs through five pruned pencil passes) + fused stencil
    (psi, V_eff, out read and written, v_{j-1}) + nonlocal box; the projection
    pass ((j+1) basis vectors + w), the CGS update ((j+1) basis vectors, v_j,
    w read, w written) and the normalisation (read + write).  Per iteration:
    the Hartree solve, the energy (n(n+1)/2 pair solves), the Ritz combination
    and the density mix."""
    P, M, c = N ** 3, 2 * N, 2 * s
    pairs = (n_occ + 1) // 2
    H = M // 2 + 1
    exch = ((n_occ + 1) * s * P + pairs * N * N * M * c            # z-forward
            + pairs * (N * N * M + N * M * M) * c                   # y-forward
            + pairs * 2 * N * M * M * c + H ** 3 * s                # x with kernel
            + pairs * (N * M * M + N * N * M) * c                   # y-inverse
            + pairs * N * N * M * c + (n_occ + 1) * s * P)          # z-inverse + epilogue
    per_step = exch + 5 * s * P + 3 * s * nl_box + 2 * s * P
    reorth = sum((j + 2) + (j + 4) for j in range(m)) * s * P
    solve = 3 * s * P + 2 * (N * N * M + N * M * M + N * M * M) * c + H ** 3 * s
    per_iter = (solve                                              # Hartree
                + (n_occ * (n_occ + 1) // 2) * solve / 2           # energy pair solves (paired)
                + (m + n_occ) * s * P                              # Ritz combination
                + (n_occ + 2) * s * P)                             # density mix
    return m * per_step + reorth + per_iter


def copy_calibration(nbytes: int = 1 << 30, reps: int = 5):
    """A plain elementwise copy of a known size (reads and writes nbytes each
    launch), larger than the 256 MB Infinity Cache so every byte crosses HBM."""
    import torch
    torch.cuda.set_device(0)
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": nbytes,
                                           "bytes_written_per_launch": nbytes,
                                           "launches": reps}}), flush=True)


def pmc_traffic(kind: int):
    """HBM bytes per launch of the probed kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor),
    or None when no summary covers this kernel kind."""
    name = PMC_SUMMARIES.get(kind)
    path = os.path.join(ROOT, "profiles", name) if name else ""
    if not name or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def cpu_baseline(N: int, h: float, n_mv: int = 4):
    """FP64 oracle on this host: time n_mv Fock matvecs (4 exchange solves each),
    a block of modified-Gram-Schmidt inner-product/axpy pairs and one Poisson
    solve at the full grid, then scale to one SCF iteration (160 matvecs,
    sum_{j<160} (j+2) MGS pairs, 11 energy solves + 1 Hartree solve)."""
    import numpy as np
    from oracle import gato_oracle as O
    g = O.Grid3D(N, h * N)
    nuc = O.water_initial_geometry(1.81, 104.5)
    el = O.water_hgh_elements()
    V = O.hgh_local_potential(nuc, el, g)
    rng = np.random.default_rng(0)
    orb = rng.standard_normal((N, N, N, 4))
    rho = O.density(orb)
    VH = O.hartree_potential(rho, g)  # also builds and caches the kernel transform
    Vnl = lambda p: O.hgh_nonlocal_apply(p, nuc, el, g)
    F = O.RHFFock(g, V, orb, VH, "dirichlet", 4, None, Vnl)
    psi = rng.standard_normal((N, N, N))
    t0 = time.perf_counter()
    for _ in range(n_mv):
        psi = F.apply(psi)
        psi = psi / np.linalg.norm(psi)
    t_mv = (time.perf_counter() - t0) / n_mv
    w = rng.standard_normal((N, N, N))
    vs = rng.standard_normal((8, N, N, N))
    t0 = time.perf_counter()
    for v in vs:
        w = w - O.inner_product(v, w, g) * v
    t_pair = (time.perf_counter() - t0) / len(vs)
    t0 = time.perf_counter()
    O.hartree_potential(rho, g)
    t_solve = time.perf_counter() - t0
    m = 160
    pairs = sum(j + 2 for j in range(m))
    t_iter = m * t_mv + pairs * t_pair + 12 * t_solve
    sample = (f"{n_mv} Fock matvecs + 8 MGS pairs + 1 Poisson solve of the 128^3 H2O workload "
              f"(numpy/scipy.fft FP64 oracle; FFTs on {O._WORKERS} threads, element-wise numpy "
              f"on one), scaled to one SCF iteration = 160 matvecs + {pairs} MGS pairs + 12 solves")
    return 1.0 / t_iter, t_iter, sample, O._WORKERS


def host_cpu() -> str:
    """CPU model and logical CPU count of this host (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs"


def uneven_batch(n_molecules: int, iters: int, streams: int) -> dict:
    """Aggregate SCF iterations/s of the seeded uneven batch on ``streams``
    concurrent streams.  iters = 0: every molecule runs FP32-mixed SCF to the
    driver's tolerance (1e-6 Ha) within the driver's iteration cap, on
    scf_rhf's default path (the H2O workload's: block Lanczos seeded with the
    occupied orbitals, stopped by its Ritz residuals, on the compressed
    exchange).  iters > 0: every molecule runs exactly ``iters``
    iterations, on one stream and then on ``streams`` streams (whose energies
    must be bit-identical).  The wall time includes each molecule's setup
    (potentials, plans, core-Hamiltonian start)."""
    from gato_amd.batch import draw_uneven_batch, eigensolver_for, scf_batch
    mols = draw_uneven_batch(n_molecules, 0)
    if iters > 0:
        log(f"uneven batch: {len(mols)} molecules x {iters} SCF iterations")
        seq = scf_batch(mols, streams=1, max_iters=iters, tol=0.0)
        log(f"  1 stream: {seq.iterations_per_s:.2f} SCF it/s ({seq.seconds:.2f} s)")
        par = scf_batch(mols, streams=streams, max_iters=iters, tol=0.0)
        log(f"  {par.streams} streams: {par.iterations_per_s:.2f} SCF it/s ({par.seconds:.2f} s)")
        same = all(a.energy == b.energy for a, b in zip(seq.points, par.points))
        return {"iters_per_s": round(par.iterations_per_s, 3), "seconds": round(par.seconds, 3),
                "streams": par.streams, "scf_iterations": par.scf_iterations,
                "sequential_iters_per_s": round(seq.iterations_per_s, 3),
                "energies_identical_to_sequential": same,
                "config": f"{len(mols)} molecules drawn with seed 0 (H2 all-electron, LiH and HCl "
                          "HGH; N in {64, 80, 96, 128}; R within 15% of equilibrium; driver box, "
                          f"mixing and Krylov defaults), {iters} FP32-mixed SCF iterations each"}
    log(f"uneven batch: {len(mols)} molecules, each to convergence")
    par = scf_batch(mols, streams=streams, tol=1e-6, eigensolver="auto")
    log(f"  {par.streams} streams: {par.iterations_per_s:.2f} SCF it/s ({par.seconds:.2f} s, "
        f"{par.scf_iterations} iterations)")
    per_kind = {}
    not_converged = [{"kind": m.kind, "N": m.N, "R": round(m.R, 4), "scf_iterations": p.n_iters}
                     for m, p in zip(mols, par.points) if not p.converged]
    for m, p in zip(mols, par.points):
        k = per_kind.setdefault(m.kind, {"molecules": 0, "converged": 0, "scf_iterations": 0,
                                         "energy_min": math.inf, "energy_max": -math.inf})
        k["molecules"] += 1
        k["converged"] += int(bool(p.converged))
        k["scf_iterations"] += p.n_iters
        k["energy_min"] = min(k["energy_min"], p.energy)
        k["energy_max"] = max(k["energy_max"], p.energy)
    return {"iters_per_s": round(par.iterations_per_s, 3), "seconds": round(par.seconds, 3),
            "streams": par.streams, "scf_iterations": par.scf_iterations,
            "all_converged": all(p.converged for p in par.points),
            "converged": sum(int(bool(p.converged)) for p in par.points),
            "not_converged": not_converged,
            "eigensolver_members": {e: sum(eigensolver_for(m, "auto") == e for m in mols)
                                    for e in sorted({eigensolver_for(m, "auto") for m in mols})},
            "per_kind": per_kind,
            "config": f"{len(mols)} molecules drawn with seed 0 (H2 all-electron, LiH and HCl "
                      "HGH; N in {64, 80, 96, 128}; R within 15% of equilibrium; driver box, "
                      "mixing, Krylov and iteration-cap defaults), each FP32-mixed to |dE| < "
                      "1e-6 Ha on scf_rhf's default path (seeded block Lanczos stopped by its "
                      "Ritz residuals, compressed exchange), as the H2O workload"}


def small_configs(reps: int = 5) -> dict:
    """The two small workloads.  Hydrogen: 64^3, side 20 bohr, softened
    -1/sqrt(r^2 + 0.1^2), 7-point Laplacian, 60 FP64 Lanczos steps from the
    normalised seed-0 Gaussian start, lowest pair; timed on the GPU (mean of
    ``reps`` runs after one warm-up) and, as its CPU baseline, the FP64 oracle
    on this host (one run).  Helium: 96^3, 16 bohr box, order 4, Z = 2
    softened Coulomb, one doubly occupied orbital, seeded random start,
    mixing 0.3, FP64 SCF to 1e-8 Ha; the second of two identical runs is
    timed (the first builds the plans)."""
    import numpy as np
    import torch
    import gato_amd as G
    N, L = 64, 20.0
    gg = G.Grid3D(N, L)
    psi0 = np.random.default_rng(0).standard_normal((N, N, N))
    psi0 /= np.sqrt(np.sum(psi0 * psi0) * gg.dV)
    H = G.Hamiltonian(gg, G.softened_coulomb(gg, Z=1.0, epsilon=0.1), order=2)
    p0 = torch.tensor(psi0, dtype=torch.float64, device="cuda")
    res = G.lanczos(H, p0, n_iters=60, n_eigenstates=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        res = G.lanczos(H, p0, n_iters=60, n_eigenstates=1)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t0) / reps
    e_gpu = float(res.eigenvalues[0])
    from oracle import gato_oracle as O   # CPU baseline of this workload
    g = O.Grid3D(N, L)
    t0 = time.perf_counter()
    ref = O.lanczos(O.Hamiltonian(g, O.softened_coulomb(g, Z=1.0, epsilon=0.1), order=2), psi0,
                    n_iters=60, n_eigenstates=1)
    cpu_s = time.perf_counter() - t0
    from gato_amd.observables import virial_ratio
    psi_g = res.states[0].contiguous()
    vr_g = float(virial_ratio(psi_g, H.V, gg, order=2))
    psi_o = ref.eigenstates[..., 0]
    vr_o = float(O.virial_ratio(psi_o, O.softened_coulomb(g, Z=1.0, epsilon=0.1), g, order=2))
    # the Ritz vector's sign is a gauge: compare |psi|
    dpsi = float(np.max(np.abs(np.abs(psi_g.cpu().numpy()) - np.abs(psi_o))))
    hydrogen = {"gpu_ms": round(gpu_s * 1e3, 3), "cpu_oracle_s": round(cpu_s, 3),
                "gpu_over_cpu": round(cpu_s / gpu_s, 1), "eigenvalue": e_gpu,
                "abs_diff_vs_oracle": abs(e_gpu - float(ref.eigenvalues[0])),
                "virial_ratio": vr_g, "virial_abs_diff_vs_oracle": abs(vr_g - vr_o),
                "state_max_abs_diff_vs_oracle": dpsi,
                "forces": "none: one nucleus at the box centre (zero by symmetry)",
                "cpu": "numpy FP64 oracle, one process (FFT-free path)",
                "config": "hydrogen 64^3 L=20, softened Coulomb a=0.1, order 2, Lanczos 60, FP64"}
    log(f"hydrogen 64^3 Lanczos 60: GPU {gpu_s * 1e3:.2f} ms, CPU oracle {cpu_s:.2f} s, "
        f"E = {e_gpu:.10f}")

    N, L = 96, 16.0
    gg = G.Grid3D(N, L)
    V = G.softened_coulomb(gg, Z=2.0)
    gen = torch.Generator(device="cuda").manual_seed(0)
    orb0 = torch.randn((N, N, N, 1), generator=gen, dtype=torch.float64, device="cuda")
    orb0 /= torch.sqrt(torch.sum(orb0 * orb0) * gg.dV)
    G.scf_rhf(V, gg, 1, initial_orbitals=orb0, max_iters=200, tol=1e-8, mixing=0.3, order=4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = G.scf_rhf(V, gg, 1, initial_orbitals=orb0, max_iters=200, tol=1e-8, mixing=0.3, order=4)
    torch.cuda.synchronize()
    he_s = time.perf_counter() - t0
    # the same solve from a second seeded start, and the one-orbital identities
    gen2 = torch.Generator(device="cuda").manual_seed(1)
    orb1 = torch.randn((N, N, N, 1), generator=gen2, dtype=torch.float64, device="cuda")
    orb1 /= torch.sqrt(torch.sum(orb1 * orb1) * gg.dV)
    r1 = G.scf_rhf(V, gg, 1, initial_orbitals=orb1, max_iters=200, tol=1e-8, mixing=0.3, order=4)
    t = G.decompose_rhf_energy(r.orbitals, V, gg, order=4)
    t1 = G.decompose_rhf_energy(r1.orbitals, V, gg, order=4)
    vir = 2 * t.kinetic / abs(t.v_ext + t.hartree + t.exchange)       # helium.py:107-112
    vir1 = 2 * t1.kinetic / abs(t1.v_ext + t1.hartree + t1.exchange)
    helium = {"seconds_to_converge": round(he_s, 3), "scf_iterations": r.n_iters,
              "scf_iters_per_s": round(r.n_iters / he_s, 2), "converged": bool(r.converged),
              "energy": float(r.energy), "orbital_energy": float(r.orbital_energies[0]),
              "virial_ratio": vir,
              "second_start": {"energy_abs_diff": abs(float(r1.energy) - float(r.energy)),
                               "orbital_energy_abs_diff": abs(float(r1.orbital_energies[0])
                                                              - float(r.orbital_energies[0])),
                               "virial_abs_diff": abs(vir1 - vir)},
              "exchange_plus_half_hartree": t.exchange + 0.5 * t.hartree,
              "forces": "none: one nucleus at the box centre (zero by symmetry)",
              "config": "helium 96^3 L=16, order 4, Z=2 softened Coulomb, 1 orbital, seeded "
                        "random start, mixing 0.3, FP64, tol 1e-8, Krylov 120"}
    log(f"helium 96^3 FP64: {r.n_iters} iterations in {he_s:.2f} s, E = {r.energy:.10f}")
    return {"hydrogen_lanczos": hydrogen, "helium_scf": helium}


def geometry_optimisation(steps: int, check_every: int, seed: int = 0) -> dict:
    """H2O geometry optimisation on 160^3 (h = 0.18, L = 28.8, HGH O/H):
    equilibrium geometry (O-H 1.81 bohr, 104.5 deg) plus a seeded uniform
    displacement of up to 0.3 bohr per coordinate, then ``steps`` Adam steps
    (lr 0.05) on Hellmann-Feynman forces, each followed by a warm-started
    FP32-mixed SCF (FP64 energies; tol 1e-7, mixing 0.5).  After every
    ``check_every``-th step the forces on all 9 coordinates are compared with
    central differences (delta = 1e-3 bohr) of the Born-Oppenheimer energy at
    fixed orbitals; the checks are timed apart from the optimisation."""
    import numpy as np
    import torch
    import gato_amd as G
    from gato_amd.geometry import Nuclei, bond_angle, bond_length
    from gato_amd.geometry_opt import optimize_geometry_alternating
    from gato_amd.physics.water import (bo_energy, hellmann_feynman_force, solve_rhf_at_geometry,
                                        water_hgh_elements, water_initial_geometry)
    grid = G.Grid3D(160, 28.8)
    el = water_hgh_elements()
    base = water_initial_geometry(1.81, 104.5)
    disp = np.random.default_rng(seed).uniform(-0.3, 0.3, (3, 3))
    nuc = Nuclei(base.positions + torch.tensor(disp, dtype=base.positions.dtype), base.charges)

    def solve_at(nuclei, prev):
        return solve_rhf_at_geometry(nuclei, el, grid, max_iters=120, tol=1e-7, mixing=0.5,
                                     precision="mixed",
                                     initial_orbitals=None if prev is None else prev.orbitals)

    def grad(p):
        return -hellmann_feynman_force(p, el, grid)

    checks, energies, scf_iters = [], [], []
    check_s = [0.0]
    d = 1e-3

    def on_step(step, p):
        energies.append(p.energy)
        scf_iters.append(p.n_iters)
        msg = f"  geom step {step + 1:3d}  E = {p.energy:+.8f}  SCF iters {p.n_iters}"
        if (step + 1) % check_every == 0:
            t0 = time.perf_counter()
            F = hellmann_feynman_force(p, el, grid).double()
            err = 0.0
            for k in range(3):
                for c in range(3):
                    Pp = p.nuclei.positions.detach().clone()
                    Pm = Pp.clone()
                    Pp[k, c] += d
                    Pm[k, c] -= d
                    Ep = float(bo_energy(Pp, p.nuclei.charges, p.orbitals, el, grid))
                    Em = float(bo_energy(Pm, p.nuclei.charges, p.orbitals, el, grid))
                    err = max(err, abs(float(F[k, c]) + (Ep - Em) / (2 * d)))
            checks.append({"step": step + 1, "max_abs_force_minus_fd": err,
                           "max_abs_force": float(F.abs().max())})
            check_s[0] += time.perf_counter() - t0
            msg += f"  |F - FD|max = {err:.2e} Ha/bohr"
        log(msg)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    final, traj = optimize_geometry_alternating(nuc, solve_at_geometry=solve_at, bo_grad=grad,
                                                geom_steps=steps, learning_rate=0.05,
                                                on_step=on_step)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0 - check_s[0]
    n_scf = sum(p.n_iters for p in traj)
    # virial ratio 2<T>/|E - <T>| and orbital energies at the final geometry
    from gato_amd.pseudopotentials import nonlocal_plan
    from gato_amd.scf import decompose_rhf_energy
    Vf = G.hgh_local_potential(final.nuclei, el, grid)
    nlf = nonlocal_plan(final.nuclei, el, grid, torch.float64)
    orb_f = final.orbitals if final.orbitals.dtype == torch.float64 else final.orbitals.double()
    tf = decompose_rhf_energy(orb_f, Vf, grid, order=4, V_nl_apply=nlf)
    virial_final = 2 * tf.kinetic / abs(tf.total + final.nuclear_repulsion - tf.kinetic)
    return {"geometry_steps": steps, "seconds": round(wall, 2),
            "geometry_steps_per_s": round(steps / wall, 4),
            "scf_iters_per_s": round(n_scf / wall, 3), "scf_iterations": n_scf,
            "start_scf_iterations": traj[0].n_iters,
            "all_converged": all(p.converged for p in traj),
            "energy_start": traj[0].energy, "energy_final": final.energy,
            "orbital_energies_final": [float(x) for x in final.orbital_energies],
            "virial_ratio_final": virial_final,
            "R_OH_final": [round(float(bond_length(final.nuclei, 0, i)), 4) for i in (1, 2)],
            "angle_final_deg": round(math.degrees(float(bond_angle(final.nuclei, 1, 0, 2))), 3),
            "force_checks": checks, "force_check_seconds": round(check_s[0], 2),
            "config": "H2O 160^3 h=0.18 HGH O/H, seeded displacement <= 0.3 bohr (seed "
                      f"{seed}), Adam lr 0.05, {steps} steps, warm-started FP32-mixed SCF "
                      "(tol 1e-7, mixing 0.5, Krylov 200), central FD delta 1e-3 on all 9 "
                      f"coordinates every {check_every} steps (timed apart)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3,
                    help="timed FP32-mixed solves of the H2O workload (each to convergence)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves before them")
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--h", type=float, default=0.2)
    ap.add_argument("--tol", type=float, default=1e-7,
                    help="SCF convergence threshold |dE| (Ha) of every H2O solve")
    ap.add_argument("--fp64-steps", type=int, default=1,
                    help="timed plain-FP64 solves of the same workload (0: skip)")
    ap.add_argument("--block-steps", type=int, default=1,
                    help="timed FP32-mixed solves with block Lanczos over the whole Krylov "
                         "dimension and the exact exchange (0: skip)")
    ap.add_argument("--refcorr-steps", type=int, default=1,
                    help="timed FP32-mixed solves whose FP64 iterations run the reference's "
                         "algorithm (scalar Lanczos, exact exchange) (0: skip)")
    ap.add_argument("--scalar-steps", type=int, default=1,
                    help="timed FP32-mixed solves with the reference's scalar Lanczos and exact "
                         "exchange in every iteration: the gain of precision alone (0: skip)")
    ap.add_argument("--no-tight-reference", action="store_true",
                    help="skip the FP64 solve to |dE| < 1e-9 that the observables' errors are "
                         "measured against")
    ap.add_argument("--batch-molecules", type=int, default=64,
                    help="independent H2/LiH/HCl molecules of the uneven batch (0: skip)")
    ap.add_argument("--batch-iters", type=int, default=0,
                    help="SCF iterations per molecule of the uneven batch (0: each molecule to "
                         "convergence at the driver's tolerance)")
    ap.add_argument("--batch-streams", type=int, default=4,
                    help="concurrent slots (streams) of the uneven batch")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the uneven batch and print its figures")
    ap.add_argument("--stencil-only", action="store_true",
                    help="only time the 16-orbital fused stencil (3 + 20 launches) and imaginary "
                         "time on H2+, and print them")
    ap.add_argument("--stencil-variant", type=int, default=-1,
                    help="with --stencil-only: force this stencil variant (see "
                         "gato_stencil_force_variant) instead of the autotuned one")
    ap.add_argument("--small-configs-only", action="store_true",
                    help="only time the hydrogen 64^3 Lanczos (GPU and CPU oracle) and the "
                         "helium 96^3 FP64 SCF")
    ap.add_argument("--no-small-configs", action="store_true",
                    help="leave the hydrogen and helium workloads out of the default run")
    ap.add_argument("--geomopt-only", action="store_true",
                    help="only run the H2O 160^3 geometry optimisation with force checks")
    ap.add_argument("--geomopt-steps", type=int, default=100)
    ap.add_argument("--geomopt-check-every", type=int, default=10)
    ap.add_argument("--geomopt-short-steps", type=int, default=100,
                    help="Adam steps of the H2O 160^3 geometry optimisation in the default run, "
                         "with force checks every --geomopt-check-every steps (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe", type=int, default=5, choices=sorted(PROBE_KINDS),
                    help="kernel kind whose launches are timed live for the roofline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (torch.mul by 1) five times: the "
                         "known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each SCF runs on one "
              "MI355X; independent molecules would be sharded across GPUs). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import gato_amd as G
    G.load_native()
    torch.cuda.set_device(0)
    if args.stencil_only:
        gbps, st_ms = stencil_bandwidth(args.N, args.h, variant=args.stencil_variant)
        it_rate, it_gbps = imag_time_rate(args.N, args.h)
        print(json.dumps({"stencil_hbm_GBps": round(gbps, 1), "ms": round(st_ms, 4),
                          "imag_time_steps_per_s": round(it_rate, 1),
                          "imag_time_algorithmic_GBps": round(it_gbps, 1)}), flush=True)
        return
    if args.small_configs_only:
        print(json.dumps(small_configs()), flush=True)
        return
    if args.geomopt_only:
        print(json.dumps({"geometry_optimisation": geometry_optimisation(
            args.geomopt_steps, args.geomopt_check_every)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batched_uneven": uneven_batch(args.batch_molecules, args.batch_iters,
                                                         args.batch_streams)}), flush=True)
        return
    N, h, tol = args.N, args.h, args.tol
    log(f"bench: H2O RHF {N}^3 h={h} (L={N * h:.2f}) on {torch.cuda.get_device_name(0)}")
    grid, nuc, el, V, nl, orb0 = build_system(N, h)

    log(f"FP32-mixed SCF solves to |dE| < {tol:g} (compressed exchange and residual-controlled "
        f"block Lanczos in the float and the FP64 iterations):")
    runs, probe = time_solves(grid, V, nl, orb0, "mixed", tol, args.warmup, args.steps,
                              args.probe, label="mixed")
    n_mixed = sum(r.n_iters for r, _ in runs)
    t_mixed = sum(dt for _, dt in runs)
    rate_mixed = n_mixed / t_mixed
    ms_per_solve = t_mixed / len(runs) * 1e3
    r_mixed = runs[-1][0]
    split_mixed = split_iterations(r_mixed)
    log(f"  -> {rate_mixed:.3f} SCF it/s over {len(runs)} solves ({ms_per_solve:.1f} ms per solve)")
    obs_mixed = h2o_observables(r_mixed, grid, V, nl, nuc, el)

    ref = None
    if not args.no_tight_reference:
        log("FP64 reference solve to |dE| < 1e-9 (the observables' errors are against it):")
        r_ref, dt_ref = solve_h2o(grid, V, nl, orb0, "fp64", 1e-9)
        ref = {"tol": 1e-9, "scf_iterations": r_ref.n_iters, "converged": bool(r_ref.converged),
               "seconds": round(dt_ref, 3),
               "observables": h2o_observables(r_ref, grid, V, nl, nuc, el)}
        log(f"  {r_ref.n_iters} iterations in {dt_ref:.3f} s, E = {r_ref.energy:+.10f}")

    def errors(obs):
        return None if ref is None else compare_observables(obs, ref["observables"])

    fp64 = None
    if args.fp64_steps > 0:
        log("plain FP64 SCF solves:")
        runs64, _ = time_solves(grid, V, nl, orb0, "fp64", tol, 0, args.fp64_steps, label="fp64")
        n64 = sum(r.n_iters for r, _ in runs64)
        t64 = sum(dt for _, dt in runs64)
        r64 = runs64[-1][0]
        obs64 = h2o_observables(r64, grid, V, nl, nuc, el)
        fp64 = {"iters_per_s": round(n64 / t64, 4), "scf_iterations": r64.n_iters,
                "seconds_to_solution": round(t64 / len(runs64), 4),
                "converged": bool(r64.converged), **split_iterations(r64),
                "observables": obs64, "errors_vs_fp64_tight": errors(obs64),
                "config": "the plain FP64 GPU path: the reference's algorithm (scalar Lanczos "
                          "with full reorthogonalisation over the whole Krylov dimension, exact "
                          "exchange in every step) in double"}
        log(f"  -> {n64 / t64:.3f} SCF it/s, {t64 / len(runs64):.3f} s to solution")

    blk = None
    if args.block_steps > 0:
        log("FP32-mixed SCF solves with block Lanczos:")
        runs_b, _ = time_solves(grid, V, nl, orb0, "mixed", tol, 0, args.block_steps,
                                eigensolver="block_lanczos", label="block")
        nb = sum(r.n_iters for r, _ in runs_b)
        tb = sum(dt for _, dt in runs_b)
        rb = runs_b[-1][0]
        obs_b = h2o_observables(rb, grid, V, nl, nuc, el)
        blk = {"iters_per_s": round(nb / tb, 4), "scf_iterations": rb.n_iters,
               "seconds_to_solution": round(tb / len(runs_b), 4), "converged": bool(rb.converged),
               **split_iterations(rb), "observables": obs_b, "errors_vs_fp64_tight": errors(obs_b),
               "config": "same workload, tolerance and precision; eigensolver block_lanczos in "
                         "every iteration: one block of the 4 occupied orbitals seeded with "
                         "them, the whole 160 Krylov vectors (40 block steps), exact exchange, "
                         "projections on the matrix cores"}
        if fp64 is not None:
            blk["errors_vs_fp64"] = compare_observables(obs_b, fp64["observables"])
            blk["time_to_solution_speedup_vs_fp64"] = round(fp64["seconds_to_solution"]
                                                            / blk["seconds_to_solution"], 3)
        log(f"  -> {nb / tb:.3f} SCF it/s, {tb / len(runs_b):.3f} s to solution")

    refc = None
    if args.refcorr_steps > 0:
        import gato_amd.scf as SCF
        log("FP32-mixed SCF solves with FP64 iterations of the reference's algorithm:")
        saved = SCF.FP64_CORRECTION
        SCF.FP64_CORRECTION = "reference"
        try:
            runs_c, _ = time_solves(grid, V, nl, orb0, "mixed", tol, 0, args.refcorr_steps,
                                    label="refcorr")
        finally:
            SCF.FP64_CORRECTION = saved
        nc = sum(r.n_iters for r, _ in runs_c)
        tc = sum(dt for _, dt in runs_c)
        rc = runs_c[-1][0]
        obs_c = h2o_observables(rc, grid, V, nl, nuc, el)
        refc = {"iters_per_s": round(nc / tc, 4), "scf_iterations": rc.n_iters,
                "seconds_to_solution": round(tc / len(runs_c), 4),
                "converged": bool(rc.converged), **split_iterations(rc), "observables": obs_c,
                "errors_vs_fp64_tight": errors(obs_c),
                "config": "same workload, tolerance and float iterations; the FP64 iterations "
                          "run the reference's algorithm (scalar Lanczos over the whole "
                          "160-vector Krylov dimension, exact exchange)"}
        if fp64 is not None:
            refc["time_to_solution_speedup_vs_fp64"] = round(fp64["seconds_to_solution"]
                                                             / refc["seconds_to_solution"], 3)
        log(f"  -> {nc / tc:.3f} SCF it/s, {tc / len(runs_c):.3f} s to solution")

    scal = None
    if args.scalar_steps > 0:
        log("FP32-mixed SCF solves with the reference's scalar Lanczos and exact exchange:")
        runs_s, _ = time_solves(grid, V, nl, orb0, "mixed", tol, 0, args.scalar_steps,
                                eigensolver="lanczos", label="scalar")
        ns = sum(r.n_iters for r, _ in runs_s)
        ts = sum(dt for _, dt in runs_s)
        rs = runs_s[-1][0]
        obs_s = h2o_observables(rs, grid, V, nl, nuc, el)
        scal = {"iters_per_s": round(ns / ts, 4), "scf_iterations": rs.n_iters,
                "seconds_to_solution": round(ts / len(runs_s), 4),
                "converged": bool(rs.converged), **split_iterations(rs), "observables": obs_s,
                "errors_vs_fp64_tight": errors(obs_s),
                "config": "same workload and tolerance; the reference's algorithm in every "
                          "iteration (scalar Lanczos over the whole 160-vector Krylov "
                          "dimension, exact exchange), FP32 iterations then FP64: the gain of "
                          "the precision pattern alone; its first solve records the Lanczos "
                          "graphs"}
        if fp64 is not None:
            scal["time_to_solution_speedup_vs_fp64"] = round(fp64["seconds_to_solution"]
                                                             / scal["seconds_to_solution"], 3)
        log(f"  -> {ns / ts:.3f} SCF it/s, {ts / len(runs_s):.3f} s to solution")

    batch = None
    if args.batch_molecules > 0:
        batch = uneven_batch(args.batch_molecules, args.batch_iters, args.batch_streams)
    small = None if args.no_small_configs else small_configs()
    geo = None
    if args.geomopt_short_steps > 0:
        log(f"H2O 160^3 geometry optimisation, {args.geomopt_short_steps} Adam steps:")
        geo = geometry_optimisation(args.geomopt_short_steps,
                                    min(args.geomopt_check_every, args.geomopt_short_steps))
    gbps, st_ms = stencil_bandwidth(N, h)
    log(f"stencil (16 orbitals, fp32): {gbps:.0f} GB/s ({st_ms:.3f} ms)")
    st_rows = stencil_batches(N, h)
    it_rate, it_gbps = imag_time_rate(N, h)
    log(f"imaginary time (H2+, fp32): {it_rate:.0f} steps/s ({it_gbps:.0f} GB/s)")

    roof = None
    if probe is not None and probe[0] > 0:
        cnt, ms, byt = probe
        avg_s = ms * 1e-3 / cnt
        achieved = byt / cnt / avg_s / 1e9
        roof = {"kernel": PROBE_KINDS[args.probe] + f"; float launches of {PROBE_SOLVES} solves "
                                                    "of the same workload run after the timed "
                                                    "ones with the probe on",
                "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                "traffic": pmc_traffic(args.probe), "algorithmic_bytes_per_launch": round(byt / cnt),
                "launches": cnt, "avg_launch_us": round(avg_s * 1e6, 2)}

    scf_roof = None
    if scal is not None and scal["fp32_iteration_ms"]:
        it_bytes = scf_iteration_bytes(N, 160, 4, 4)
        ms32 = scal["fp32_iteration_ms"]
        scf_roof = {"iteration": "one FP32 iteration of the reference's algorithm (scalar "
                                 "Lanczos, 160 steps, exact exchange): the scalar leg",
                    "algorithmic_GB_per_fp32_iteration": round(it_bytes / 1e9, 1),
                    "floor_ms_at_8TBps": round(it_bytes / (HBM_PEAK_GBPS * 1e9) * 1e3, 2),
                    "fp32_iteration_ms": ms32,
                    "frac": round(it_bytes / (HBM_PEAK_GBPS * 1e9) / (ms32 * 1e-3), 4)}

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU oracle sample ...")
        os.environ.setdefault("OMP_NUM_THREADS", str(os.cpu_count()))
        c_rate, c_iter, sample, threads = cpu_baseline(N, h)
        cpu = {"value": round(c_rate, 6), "unit": "SCF iters/s", "cores": threads,
               "kind": "port", "sample": sample, "s_per_iter": round(c_iter, 2),
               "host": host_cpu()}
        log(f"  CPU oracle: {c_rate:.5f} SCF it/s ({c_iter:.1f} s/it)")

    ph = [p for p in r_mixed.precision_history if p != "fp64-eval"]
    mixed = {"seconds_to_solution": round(ms_per_solve * 1e-3, 4),
             "scf_iterations": r_mixed.n_iters, "converged": bool(r_mixed.converged),
             **split_mixed, "observables": obs_mixed,
             "errors_vs_fp64_tight": errors(obs_mixed),
             "solves_timed": len(runs),
             "iterations_per_solve": [r.n_iters for r, _ in runs],
             "krylov_vectors_fp32_iterations": [k for k, p in zip(r_mixed.krylov_history, ph)
                                                if p == "fp32"],
             "krylov_vectors_fp64_iterations": [k for k, p in zip(r_mixed.krylov_history, ph)
                                                if p == "fp64"],
             "config": "FP32 iterations: fields, stencils, FFTs and the Krylov basis in float, "
                       "every inner product, the Lanczos coefficients and the eigensolves in "
                       "double; the eigensolver is block Lanczos seeded with the current "
                       "orbitals, stopped once its Ritz residuals are below "
                       "min(1, max(1e-4, 0.05 sqrt|dE|)), on the Fock operator with the "
                       "compressed exchange of the current orbitals (built from their 16 "
                       "exchange solves, exact on them).  Then FP64 iterations of the same "
                       "scheme, with Ritz residuals below 0.05 sqrt(tol), run until two "
                       "consecutive ones differ by less than tol"}
    if fp64 is not None:
        mixed["errors_vs_fp64"] = compare_observables(obs_mixed, fp64["observables"])
        mixed["time_to_solution_speedup_vs_fp64"] = round(fp64["seconds_to_solution"]
                                                          / mixed["seconds_to_solution"], 3)
        mixed["iters_per_s_ratio_vs_fp64"] = round(rate_mixed / fp64["iters_per_s"], 3)
    line = {
        "metric": METRIC,
        "value": round(rate_mixed, 4),
        "unit": "SCF iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_solve, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded random starting orbitals; HGH H2O geometry O-H 1.81 bohr, "
                "104.5 deg)",
        "config": {"workload": f"H2O RHF SCF solve to convergence (|dE| < {tol:g} Ha) from seeded "
                               f"random orbitals, {N}^3 grid h={h} bohr, HGH O/H, 4 occupied, "
                               "order-4 stencil, Krylov 160, mixing 0.5, FP32-mixed (FP32 "
                               "iterations with compressed exchange and residual-controlled "
                               "block Lanczos, then FP64 iterations of the same scheme to "
                               "tol); one step = one solve, value = "
                               "SCF iterations / wall seconds over the timed solves",
                   "N": N, "h": h, "n_occ": 4, "krylov": 160, "tol": tol,
                   "parallelism": "single GPU"},
        "fp32_mixed": mixed,
        "fp64": fp64,
        "fp32_mixed_reference_fp64": refc,
        "fp32_mixed_scalar_lanczos": scal,
        "fp64_tight_reference": ref,
        "stencil_hbm_GBps": round(gbps, 1),
        "stencil_batches": st_rows,
        "block_lanczos": blk,
        "batched_uneven": batch,
        "small_configs": small,
        "geometry_optimisation": geo,
        "imag_time_h2plus": {"steps_per_s": round(it_rate, 1), "algorithmic_GBps": round(it_gbps, 1),
                             "config": f"H2+ R=2 bohr, {N}^3 h={h}, order 4, fp32, Euler step + "
                                       "renormalisation"},
        "roofline": roof,
        "scf_roofline": scf_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
