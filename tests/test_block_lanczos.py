"""Block Lanczos: the oracle against dense eigensolves, the degenerate level a
single start vector cannot reach (solvers/lanczos.py:83-107), and the
library's host-side eigensolvers (Householder + QL, block Rayleigh-Ritz)
against numpy.  No GPU: the host functions of libgato_amd.so run on the CPU."""
import ctypes as C

import numpy as np
import pytest

from oracle import gato_oracle as O


def dense_matrix(H):
    """Assemble H column by column (brute force, small grids only)."""
    shape = H.grid.shape
    n = int(np.prod(shape))
    M = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        M[:, j] = H.apply(e.reshape(shape)).ravel()
        e[j] = 0.0
    return 0.5 * (M + M.T)


def test_block_lanczos_matches_dense_eigh():
    rng = np.random.default_rng(5)
    g = O.Grid3D(6, 6.0)
    H = O.Hamiltonian(g, 0.5 * rng.standard_normal(g.shape), order=4)
    ref = np.linalg.eigh(dense_matrix(H))[0]
    X0 = rng.standard_normal(g.shape + (4,))
    res = O.block_lanczos(H, X0, n_steps=40, n_eigenstates=4)
    assert np.max(np.abs(res.eigenvalues - ref[:4])) < 1e-8
    # Ritz vectors are eigenvectors of H with the grid's inner product
    for k in range(4):
        x = res.eigenstates[..., k]
        r = H.apply(x) - res.eigenvalues[k] * x
        assert np.sqrt(O.norm_sq(r, g)) < 1e-6
        assert abs(O.norm_sq(x, g) - 1.0) < 1e-10
    # T blocks: A symmetric, B upper triangular with a positive diagonal
    for A, B in zip(res.alphas, res.betas):
        assert np.allclose(A, A.T, atol=1e-14)
        assert np.allclose(B, np.triu(B)) and np.all(np.diag(B) > 0)


def test_block_lanczos_resolves_degenerate_level():
    """The isotropic oscillator's first excited level is exactly threefold on a
    cubic grid.  A single Krylov sequence holds one direction of a degenerate
    eigenspace; its partners only enter through noise and round-off, slowly
    (lanczos.py:83-107).  At the same Krylov dimension (120 vectors) scalar
    Lanczos still misses part of the triplet, while block Lanczos seeded with
    four vectors returns all of it."""
    g = O.Grid3D(10, 8.0)
    H = O.Hamiltonian(g, O.harmonic_oscillator(g), order=4)
    ref = np.linalg.eigh(dense_matrix(H))[0]
    assert ref[3] - ref[1] < 1e-10 and ref[4] - ref[3] > 0.1   # the triplet
    X, Y, Z = g.coords()
    sym = np.exp(-0.5 * (X * X + Y * Y + Z * Z))
    rng = np.random.default_rng(2)
    scalar = O.lanczos(H, sym + 0.01 * rng.standard_normal(g.shape), n_iters=120,
                       n_eigenstates=4)
    assert abs(scalar.eigenvalues[0] - ref[0]) < 1e-9
    assert scalar.eigenvalues[3] > ref[3] + 0.1                # a copy is missing
    X0 = np.stack([sym] + [sym * rng.standard_normal(g.shape) for _ in range(3)], axis=-1)
    block = O.block_lanczos(H, X0, n_steps=30, n_eigenstates=4)
    assert np.max(np.abs(block.eigenvalues - ref[:4])) < 1e-8


@pytest.fixture(scope="module")
def lib():
    from gato_amd import _lib as L
    return L.lib()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.mark.parametrize("m", [1, 2, 7, 50])
def test_sym_eigh_matches_numpy(lib, m):
    rng = np.random.default_rng(m)
    M = rng.standard_normal((m, m))
    M = np.ascontiguousarray(M + M.T)
    M[0, :] = M[:, 0] = 0.0  # an already-reduced column
    M[0, 0] = 3.0
    K = min(m, 5)
    ev = np.empty(K)
    V = np.empty((m, K))
    assert lib.gato_sym_eigh(m, _dp(M), K, _dp(ev), _dp(V)) == 0
    w, U = np.linalg.eigh(M)
    assert np.max(np.abs(ev - w[:K])) < 1e-12 * max(1.0, np.max(np.abs(w)))
    assert np.allclose(M @ V, V * ev, atol=1e-11)
    assert np.allclose(V.T @ V, np.eye(K), atol=1e-12)
    for k in range(K):  # gauge: largest-magnitude component positive
        assert V[np.argmax(np.abs(V[:, k])), k] > 0


def test_block_ritz_matches_dense_and_drops_columns(lib):
    rng = np.random.default_rng(9)
    nb, s, K = 3, 5, 4
    m = nb * s
    A = np.stack([(lambda a: a + a.T)(rng.standard_normal((nb, nb))) for _ in range(s)])
    B = np.stack([np.triu(rng.standard_normal((nb, nb))) + 2 * np.eye(nb) for _ in range(s)])
    T = np.zeros((m, m))
    for i in range(s):
        T[i * nb:(i + 1) * nb, i * nb:(i + 1) * nb] = A[i]
        if i + 1 < s:
            T[(i + 1) * nb:(i + 2) * nb, i * nb:(i + 1) * nb] = B[i]
            T[i * nb:(i + 1) * nb, (i + 1) * nb:(i + 2) * nb] = B[i].T
    drop = np.zeros((s + 1, nb), dtype=np.int32)
    ev = np.empty(K)
    S = np.empty((m, K))
    ip = C.POINTER(C.c_int)
    assert lib.gato_block_ritz(nb, s, _dp(A), _dp(B), drop.ctypes.data_as(ip), K, _dp(ev),
                               _dp(S)) == 0
    assert np.max(np.abs(ev - np.linalg.eigh(T)[0][:K])) < 1e-12
    assert np.allclose(T @ S, S * ev, atol=1e-11)
    # a dropped column (a zero basis vector) leaves T
    drop[2, 1] = 1
    keep = [j for j in range(m) if j != 2 * nb + 1]
    assert lib.gato_block_ritz(nb, s, _dp(A), _dp(B), drop.ctypes.data_as(ip), K, _dp(ev),
                               _dp(S)) == 0
    Tk = T[np.ix_(keep, keep)]
    assert np.max(np.abs(ev - np.linalg.eigh(Tk)[0][:K])) < 1e-12
    assert np.all(S[2 * nb + 1] == 0.0)


@pytest.mark.parametrize("nb,s,tri", [(4, 40, True), (8, 20, True), (4, 25, False), (1, 60, True)])
def test_block_ritz_band_solver_matches_numpy(lib, nb, s, tri):
    """The band reduction behind gato_block_ritz on the sizes of the H2O runs
    (up to 160 vectors) and with a full (not triangular) coupling block, whose
    band is 2 nb - 1 wide instead of nb."""
    rng = np.random.default_rng(nb * 100 + s)
    m = nb * s
    K = min(8, m)
    A = np.stack([(lambda a: a + a.T)(rng.standard_normal((nb, nb))) for _ in range(s)])
    B = rng.standard_normal((s, nb, nb))
    if tri:
        B = np.triu(B)
    T = np.zeros((m, m))
    for i in range(s):
        T[i * nb:(i + 1) * nb, i * nb:(i + 1) * nb] = A[i]
        if i + 1 < s:
            T[(i + 1) * nb:(i + 2) * nb, i * nb:(i + 1) * nb] = B[i]
            T[i * nb:(i + 1) * nb, (i + 1) * nb:(i + 2) * nb] = B[i].T
    drop = np.zeros((s + 1, nb), dtype=np.int32)
    ev = np.empty(K)
    S = np.empty((m, K))
    assert lib.gato_block_ritz(nb, s, _dp(np.ascontiguousarray(A)), _dp(np.ascontiguousarray(B)),
                               drop.ctypes.data_as(C.POINTER(C.c_int)), K, _dp(ev), _dp(S)) == 0
    ref = np.linalg.eigvalsh(T)[:K]
    assert np.max(np.abs(ev - ref)) < 1e-11 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(T @ S - S * ev)) < 1e-9
    assert np.max(np.abs(S.T @ S - np.eye(K))) < 1e-8
