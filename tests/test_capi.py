"""The C-ABI library loads without a GPU, exports every symbol that
include/gato_amd.h declares, and its host-side routines are correct."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "gato_amd.h").read_text()
    return sorted(set(re.findall(r"\b(gato_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from gato_amd import _lib
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding covers the same set
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)
    assert lib.gato_abi_version() == 1


@pytest.mark.parametrize("m,K", [(1, 1), (2, 2), (7, 3), (60, 4), (160, 4), (200, 12)])
def test_tridiag_eigh_matches_dense_eigh(m, K):
    """Host eigensolver (Sturm bisection + inverse iteration) vs numpy.linalg.eigh (the reference calls
    jnp.linalg.eigh on the same tridiagonal, lanczos.py:176-180)."""
    from gato_amd.solvers.lanczos import _tridiag
    rng = np.random.default_rng(m)
    a = rng.standard_normal(m) * 3
    b = np.abs(rng.standard_normal(m)) + 1e-3
    ev, S = _tridiag(a, b, K)
    T = np.diag(a) + np.diag(b[: m - 1], 1) + np.diag(b[: m - 1], -1)
    w, U = np.linalg.eigh(T)
    assert np.max(np.abs(ev - w[:K])) < 1e-12 * max(1, np.max(np.abs(w)))
    for k in range(K):
        s = np.sign(np.dot(S[:, k], U[:, k]))
        assert np.max(np.abs(s * S[:, k] - U[:, k])) < 1e-9


def test_tridiag_eigh_clustered_spectrum():
    from gato_amd.solvers.lanczos import _tridiag
    m = 40
    a = np.ones(m)
    b = np.full(m, 1e-9)
    ev, S = _tridiag(a, b, 5)
    T = np.diag(a) + np.diag(b[: m - 1], 1) + np.diag(b[: m - 1], -1)
    w = np.linalg.eigvalsh(T)
    assert np.max(np.abs(ev - w[:5])) < 1e-14
    assert np.max(np.abs(S.T @ S - np.eye(5))) < 1e-12


def test_bad_arguments_map_to_value_error():
    from gato_amd import _lib
    import ctypes as C
    lib = _lib.load()
    st = lib.gato_tridiag_eigh(0, None, None, 1, None, None)
    with pytest.raises(ValueError):
        _lib.check(st)
    st = lib.gato_stencil_apply(0, 8, 1, 0.1, 3, 0, 1.0, None, None, 0, None, None, None, None)
    with pytest.raises(ValueError, match="order"):
        _lib.check(st)
    # imaginary time: argument checks run before any device work
    fake_a, fake_b, fake_ws = C.c_void_p(256), C.c_void_p(512), C.c_void_p(1024)
    for args, msg in [((0, 8, 0.1, 3, 0, 0.01, None, fake_a, fake_b, 1, 0.001, None, fake_ws, None),
                       "order"),
                      ((0, 8, 0.1, 4, 0, 0.01, None, fake_a, fake_a, 1, 0.001, None, fake_ws, None),
                       "bad arguments"),
                      ((0, 8, 0.1, 4, 0, 0.01, None, fake_a, fake_b, -1, 0.001, None, fake_ws, None),
                       "bad arguments"),
                      ((2, 8, 0.1, 4, 0, 0.01, None, fake_a, fake_b, 1, 0.001, None, fake_ws, None),
                       "dtype")]:
        with pytest.raises(ValueError, match=msg):
            _lib.check(lib.gato_imag_time_steps(*args))
