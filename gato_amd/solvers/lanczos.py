"""Matrix-free Lanczos eigensolver with full reorthogonalisation
(reference: solvers/lanczos.py:1-206).

Operators built by this package (``Hamiltonian``, ``RHFFock``) run the whole
m-step loop natively (gato_lanczos: one stream, no host round trip until the
tridiagonal eigensolve).  Any other duck-typed operator with ``.grid`` and
``.apply(psi)`` (the reference's protocol, lanczos.py:121) runs a Python loop
around its ``apply`` that still uses the native dot / Gram-Schmidt kernels.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _lib as L
from .. import _native as nat


@dataclass
class LanczosResult:
    """lanczos.py:41-46; eigenstates is an (N, N, N, K) view of a (K, N, N, N) stack."""

    eigenvalues: torch.Tensor   # (K,) float64, ascending
    eigenstates: torch.Tensor   # (N, N, N, K)
    n_iters: int
    converged: bool             # True if any beta fell below tol (lanczos.py:199)
    alphas: np.ndarray = field(default=None, repr=False)
    betas: np.ndarray = field(default=None, repr=False)
    residual: float | None = None  # largest Ritz residual of an adaptive run's last check

    @property
    def states(self) -> torch.Tensor:
        """The Ritz vectors as a contiguous (K, N, N, N) stack."""
        return self.eigenstates.permute(3, 0, 1, 2)


def default_krylov_dim(N: int, n_eigenstates: int = 1) -> int:
    """max(24 n + 24, floor(5N/4)) (lanczos.py:49-107)."""
    return max(24 * n_eigenstates + 24, (5 * N) // 4)


def _tridiag(alphas: np.ndarray, betas: np.ndarray, K: int):
    m = len(alphas)
    ev = (C.c_double * K)()
    S = (C.c_double * (m * K))()
    a = (C.c_double * m)(*alphas.tolist())
    b = (C.c_double * max(m, 1))(*betas.tolist())
    L.check(L.lib().gato_tridiag_eigh(m, a, b, K, ev, S))
    return np.array(ev[:]), np.array(S[:]).reshape(m, K)


def lanczos(H, psi0: torch.Tensor, n_iters: int = 80, n_eigenstates: int = 4,
            tol: float = 1e-10) -> LanczosResult:
    """Lowest ``n_eigenstates`` eigenpairs of H after ``n_iters`` Lanczos steps
    (lanczos.py:110-206).  All ``n_iters`` steps always run."""
    grid = H.grid
    psi0 = nat.contig(psi0, "psi0")
    m = int(n_iters)
    K = min(int(n_eigenstates), m)
    if m < 1 or K < 1:
        raise ValueError("n_iters and n_eigenstates must be >= 1")
    native = getattr(H, "_native_fock", None)
    if native is not None:
        return _lanczos_native(H, native, psi0, m, K, tol)
    return _lanczos_generic(H, psi0, m, K, tol)


def _lanczos_native(H, native_ctx, psi0, m, K, tol) -> LanczosResult:
    grid = H.grid
    n3 = grid.N ** 3
    with native_ctx(psi0.dtype) as fock:
        V = torch.empty((m + 1, *grid.shape), dtype=psi0.dtype, device=psi0.device)
        scratch = torch.empty(int(L.lib().gato_lanczos_scratch_bytes(m, K)), dtype=torch.uint8,
                              device=psi0.device)
        states = torch.empty((K, *grid.shape), dtype=psi0.dtype, device=psi0.device)
        alphas = (C.c_double * m)()
        betas = (C.c_double * m)()
        evals = (C.c_double * K)()
        ws = L.workspace(psi0.device)
        L.check(L.lib().gato_lanczos(fock, psi0.data_ptr(), m, K, float(grid.dV), V.data_ptr(),
                                     scratch.data_ptr(), alphas, betas, evals, states.data_ptr(),
                                     ws.data_ptr(), L.stream_ptr()))
        del V
    a = np.array(alphas[:])
    b = np.array(betas[:])
    converged = bool(np.any(b[: max(m - 1, 1)] < tol))
    ev = torch.tensor(evals[:], dtype=torch.float64)
    return LanczosResult(ev, states.permute(1, 2, 3, 0), m, converged, a, b)


def _lanczos_generic(H, psi0, m, K, tol) -> LanczosResult:
    grid = H.grid
    n3 = grid.N ** 3
    dt = psi0.dtype
    dev = psi0.device
    lib = L.lib()
    code = L.dtype_code(dt)
    ws = L.workspace(dev).data_ptr()
    sp = L.stream_ptr()
    V = torch.empty((m + 1, *grid.shape), dtype=dt, device=dev)
    coef = torch.empty(m + 1, dtype=torch.float64, device=dev)
    alpha = torch.empty(m, dtype=torch.float64, device=dev)
    beta = torch.empty(m + 1, dtype=torch.float64, device=dev)
    spare = beta[m:].data_ptr()
    # v0 = psi0 / ||psi0||  (lanczos.py:134)
    L.check(lib.gato_normalize(code, n3, psi0.data_ptr(), grid.dV, spare, V[0].data_ptr(), ws, sp))
    for j in range(m):
        w = H.apply(V[j]).to(dt).contiguous()
        vj = V[j].data_ptr()
        if j > 0:  # w -= beta_{j-1} v_{j-1} (lanczos.py:146-148); norm goes to a spare slot
            L.check(lib.gato_cgs_update(code, n3, 0, None, 0, None, V[j - 1].data_ptr(),
                                        beta[j - 1:].data_ptr(), w.data_ptr(), w.data_ptr(),
                                        grid.dV, spare, ws, sp))
        aj = alpha[j:].data_ptr()
        # alpha_j = <v_j, w> (lanczos.py:150) and c_k = <v_k, w - alpha_j v_j> in one pass
        L.check(lib.gato_cgs_coefficients(code, n3, j + 1, V.data_ptr(), n3, w.data_ptr(),
                                          grid.dV, coef.data_ptr(), aj, ws, sp))
        # w -= alpha_j v_j + sum c_k v_k; beta_j = ||w|| (lanczos.py:151-163)
        L.check(lib.gato_cgs_update(code, n3, j + 1, V.data_ptr(), n3, coef.data_ptr(), vj, aj,
                                    w.data_ptr(), V[j + 1].data_ptr(), grid.dV,
                                    beta[j:].data_ptr(), ws, sp))
        L.check(lib.gato_scale_by_inverse(code, n3, V[j + 1].data_ptr(), beta[j:].data_ptr(),
                                          V[j + 1].data_ptr(), sp))
    a = alpha.cpu().numpy()
    b = beta[:m].cpu().numpy()
    ev, S = _tridiag(a, b, K)
    Sd = torch.tensor(S, dtype=torch.float64, device=dev).contiguous()
    states = torch.empty((K, *grid.shape), dtype=dt, device=dev)
    L.check(lib.gato_ritz_combine(code, n3, m, K, V.data_ptr(), n3, Sd.data_ptr(),
                                  states.data_ptr(), n3, sp))
    converged = bool(np.any(b[: max(m - 1, 1)] < tol))
    return LanczosResult(torch.tensor(ev), states.permute(1, 2, 3, 0), m, converged, a, b)


_GRAPH_STREAMS: dict = {}
# Recording a graph and destroying one exclude each other: HIP refuses to
# destroy a graph while any stream records (the slots of a batch record and
# finish molecules concurrently).  A dropped graph is parked in _RETIRED and
# destroyed by the next recording thread before it records, under the lock.
_CAPTURE_LOCK = threading.Lock()
_RETIRED: list = []
_RETIRED_LOCK = threading.Lock()


def _release_retired() -> None:
    """Destroy the parked graphs (call with _CAPTURE_LOCK held)."""
    with _RETIRED_LOCK:
        held = _RETIRED[:]
        _RETIRED.clear()
    del held


def _graph_stream(device: torch.device) -> torch.cuda.Stream:
    """The side stream on which graphs for work issued on the current stream
    are recorded: one per (device, current stream), kept for the process, so
    the Poisson plans and workspace created for it once serve every later
    recording.  Replays run on the current stream itself (stream-ordered with
    the rest of its work); the two streams map one to one, so replays from
    different streams never share a plan or workspace."""
    key = (L.device_index(device), L.stream_ptr(device))
    s = _GRAPH_STREAMS.get(key)
    if s is None:
        s = torch.cuda.Stream(device=device)
        _GRAPH_STREAMS[key] = s
    return s


class LanczosGraph:
    """Native Lanczos (gato_lanczos_steps) on fixed buffers, recorded into a
    HIP graph and replayed: the m-step loop of lanczos.py:134-167 for an
    operator whose inputs change in place between runs (the SCF loop's Fock
    operator: the effective potential and the exchange orbitals).

    The caller writes ``veff``, ``orb`` and ``seed`` and then calls
    ``run()``.  The first run executes the steps eagerly on the recording
    stream (building its plans and workspace, tuning the stencil), the second
    records them into a graph and replays it, later runs only replay.  The
    kernels, their order and their arguments are the same in every run, so
    every run gives the eager native result bit for bit (the stencil variant
    chosen under recording differs from the tuned one only in speed).  The
    tridiagonal eigensolve (host) and the Ritz combination follow each run on
    the current stream."""

    def __init__(self, grid, dtype: torch.dtype, n_orb: int, m: int, K: int, device,
                 order: int, boundary: str, epsilon, nl):
        from .poisson import get_plan
        self.grid, self.dtype, self.m, self.K = grid, dtype, int(m), int(K)
        self.order, self.boundary = order, boundary
        dev = torch.device(device)
        self.device = dev
        self.veff = torch.zeros(grid.shape, dtype=dtype, device=dev)
        self.orb = torch.zeros((n_orb, *grid.shape), dtype=dtype, device=dev)
        self.seed = torch.zeros(grid.shape, dtype=dtype, device=dev)
        self.V = torch.empty((m + 1, *grid.shape), dtype=dtype, device=dev)
        self.scratch = torch.zeros(int(L.lib().gato_lanczos_scratch_bytes(m, K)),
                                   dtype=torch.uint8, device=dev)
        self.nl = nl
        self.stream = _graph_stream(dev)
        self.graph = None
        self.runs = 0
        with torch.cuda.stream(self.stream):
            self.plan = get_plan(grid, epsilon, dtype, dev)
            self.ws = L.workspace(dev)
        self._fock = C.c_void_p()
        n3 = grid.N ** 3
        L.check(L.lib().gato_fock_create(
            L.dtype_code(dtype), grid.N, grid.h, nat.check_order(order), nat.boundary_code(boundary),
            self.veff.data_ptr(), self.plan.handle, self.orb.data_ptr(), n3, n_orb,
            nl.handle if nl is not None else None, C.byref(self._fock)))

    def _steps(self) -> None:
        L.check(L.lib().gato_lanczos_steps(self._fock, self.seed.data_ptr(), self.m, self.K,
                                           float(self.grid.dV), self.V.data_ptr(),
                                           self.scratch.data_ptr(), self.ws.data_ptr(),
                                           L.stream_ptr()))

    def run(self, tol: float = 1e-10) -> LanczosResult:
        cur = torch.cuda.current_stream(self.device)
        if self.runs < 2:
            # the recording stream waits for the inputs written on the current one
            self.stream.wait_stream(cur)
            if self.runs == 0:
                with torch.cuda.stream(self.stream):
                    self._steps()
            else:
                with _CAPTURE_LOCK:
                    _release_retired()
                    g = torch.cuda.CUDAGraph()
                    # thread_local: other host threads (the slots of a batch)
                    # may allocate or synchronise while this one records
                    with torch.cuda.graph(g, stream=self.stream,
                                          capture_error_mode="thread_local"):
                        self._steps()
                    self.graph = g
            cur.wait_stream(self.stream)
        if self.graph is not None:
            self.graph.replay()
        self.runs += 1
        m, K = self.m, self.K
        ab = self.scratch.view(torch.float64)[m + 1:3 * m + 1].cpu().numpy()
        a, b = ab[:m].copy(), ab[m:].copy()
        ev, S = _tridiag(a, b, K)
        Sd = torch.tensor(S, dtype=torch.float64, device=self.device).contiguous()
        n3 = self.grid.N ** 3
        states = torch.empty((K, *self.grid.shape), dtype=self.dtype, device=self.device)
        L.check(L.lib().gato_ritz_combine(L.dtype_code(self.dtype), n3, m, K, self.V.data_ptr(), n3,
                                          Sd.data_ptr(), states.data_ptr(), n3, L.stream_ptr()))
        converged = bool(np.any(b[: max(m - 1, 1)] < tol))
        return LanczosResult(torch.tensor(ev), states.permute(1, 2, 3, 0), m, converged, a, b)

    def close(self) -> None:
        if self._fock:
            L.lib().gato_fock_destroy(self._fock)  # host memory only
            self._fock = C.c_void_p()
        if self.graph is not None:
            with _RETIRED_LOCK:
                _RETIRED.append(self.graph)
            self.graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
