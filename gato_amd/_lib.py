"""ctypes binding of the gato_amd C ABI (include/gato_amd.h).

The shared library ``libgato_amd.so`` is built in-tree by ``gato_amd/build.py``.
There is no CPU fallback: every function here runs HIP kernels on the current
torch CUDA (ROCm) device and stream, and loading fails loudly if the library
is missing.  Status codes map to the exceptions the reference raises for the
same conditions (ValueError, KeyError) and RuntimeError for HIP failures.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import threading
from pathlib import Path

import torch

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libgato_amd.so"

GATO_OK = 0
GATO_ERR_VALUE = 1
GATO_ERR_KEY = 2
GATO_ERR_HIP = 3
GATO_ERR_UNSUPPORTED = 4

F32 = 0
F64 = 1
DIRICHLET = 0
PERIODIC = 1

CENTER_SOFT_COULOMB = 0
CENTER_HGH_LOCAL = 1
CENTER_HARMONIC = 2
CENTER_SLATER_1S = 3
CENTER_PARAMS = 10

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i = C.c_int
_l = C.c_long
_d = C.c_double

# name -> (restype, argtypes); every symbol declared in include/gato_amd.h
SIGNATURES = {
    "gato_last_error": (C.c_char_p, []),
    "gato_abi_version": (_i, []),
    "gato_reduce_workspace_bytes": (C.c_size_t, []),
    "gato_probe_reset": (None, [_i]),
    "gato_set_autotune": (None, [_i]),
    "gato_stencil_force_variant": (None, [_i]),
    "gato_poisson_force_pencils": (None, [_i]),
    "gato_probe_read": (_i, [C.POINTER(_l), C.POINTER(_d), C.POINTER(_d)]),
    "gato_stencil_apply": (_i, [_i, _i, _i, _d, _i, _i, _d, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "gato_stencil_normalize": (_i, [_i, _i, _d, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp,
                                    _vp]),
    "gato_imag_time_steps": (_i, [_i, _i, _d, _i, _i, _d, _vp, _vp, _vp, _i, _d, _vp, _vp, _vp]),
    "gato_dot_batched": (_i, [_i, _l, _i, _vp, _l, _vp, _l, _vp, _vp, _vp, _d, _vp, _vp, _vp]),
    "gato_dot_batched_noalloc": (_i, [_i, _l, _i, _vp, _l, _vp, _l, _vp, _vp, _vp, _d, _vp, _vp]),
    "gato_cgs_update": (_i, [_i, _l, _i, _vp, _l, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp]),
    "gato_cgs_coefficients": (_i, [_i, _l, _i, _vp, _l, _vp, _d, _vp, _vp, _vp, _vp]),
    "gato_scale_by_inverse": (_i, [_i, _l, _vp, _vp, _vp, _vp]),
    "gato_normalize": (_i, [_i, _l, _vp, _d, _vp, _vp, _vp, _vp]),
    "gato_ritz_combine": (_i, [_i, _l, _i, _i, _vp, _l, _vp, _vp, _l, _vp]),
    "gato_tridiag_eigh": (_i, [_i, _dp, _dp, _i, _dp, _dp]),
    "gato_density_mix": (_i, [_i, _l, _i, _vp, _l, _d, _vp, _vp, _vp]),
    "gato_axpby": (_i, [_i, _l, _d, _vp, _d, _vp, _vp]),
    "gato_convert": (_i, [_i, _vp, _i, _vp, _l, _vp]),
    "gato_local_potential": (_i, [_i, _i, _d, _i, _dp, _i, _vp, _vp]),
    "gato_local_potential_grad": (_i, [_i, _i, _d, _i, _dp, _vp, _vp, _vp, _vp]),
    "gato_poisson_plan_create": (_i, [_i, _i, _d, _d, _i, C.POINTER(_vp), _vp]),
    "gato_poisson_plan_destroy": (_i, [_vp]),
    "gato_hartree_potential": (_i, [_vp, _i, _vp, _l, _vp, _l, _vp]),
    "gato_hartree_energy": (_i, [_vp, _i, _vp, _l, _vp, _vp, _vp]),
    "gato_exchange_apply": (_i, [_vp, _vp, _vp, _l, _i, _d, _i, _vp, _vp]),
    "gato_exchange_pair_integrals": (_i, [_vp, _vp, _l, _i, _vp, _vp, _vp]),
    "gato_nl_plan_create": (_i, [_i, _i, _d, _i, _dp, _i, _ip, _ip, _ip, _ip, _dp, _dp, _d,
                                 C.POINTER(_vp), _vp]),
    "gato_nl_plan_destroy": (_i, [_vp]),
    "gato_nl_plan_points": (_l, [_vp]),
    "gato_nl_apply": (_i, [_vp, _i, _vp, _l, _vp, _l, _vp, _vp]),
    "gato_nl_expectation": (_i, [_vp, _i, _vp, _l, _vp, _vp, _vp]),
    "gato_nl_grad": (_i, [_vp, _i, _vp, _l, _d, _vp, _vp, _vp]),
    "gato_fock_create": (_i, [_i, _i, _d, _i, _i, _vp, _vp, _vp, _l, _i, _vp, C.POINTER(_vp)]),
    "gato_fock_destroy": (_i, [_vp]),
    "gato_fock_apply": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gato_lanczos_scratch_bytes": (C.c_size_t, [_i, _i]),
    "gato_lanczos": (_i, [_vp, _vp, _i, _i, _d, _vp, _vp, _dp, _dp, _dp, _vp, _vp, _vp]),
    "gato_lanczos_steps": (_i, [_vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp]),
    "gato_block_dot": (_i, [_i, _l, _i, _vp, _l, _vp, _l, _i, _d, _vp, _vp, _vp]),
    "gato_block_update": (_i, [_i, _l, _i, _vp, _l, _vp, _vp, _l, _i, _vp, _l, _d, _vp, _vp, _vp]),
    "gato_block_step_scratch_bytes": (C.c_size_t, [_i, _i]),
    "gato_block_orthonormalize": (_i, [_i, _l, _i, _vp, _l, _vp, _l, _d, _d, _vp, _vp, _vp, _vp,
                                       _vp]),
    "gato_block_lanczos_step": (_i, [_i, _l, _i, _i, _vp, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp]),
    "gato_block_ritz": (_i, [_i, _i, _dp, _dp, _ip, _i, _dp, _dp]),
    "gato_sym_eigh": (_i, [_i, _dp, _i, _dp, _dp]),
    "gato_block_scratch_bytes": (C.c_size_t, [_i, _i]),
    "gato_block_lanczos": (_i, [_vp, _vp, _l, _i, _i, _i, _d, _d, _vp, _vp, _dp, _dp, _ip, _dp,
                                _vp, _vp, _vp]),
    "gato_ace_build": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "gato_fock_set_ace": (_i, [_vp, _vp, _i, _vp]),
    "gato_block_lanczos_adaptive": (_i, [_vp, _vp, _l, _i, _i, _i, _d, _i, _d, _d, _vp, _vp, _dp,
                                         _dp, _ip, _dp, _vp, _ip, _dp, _vp, _vp]),
}

_lib = None
_lock = threading.Lock()


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libgato_amd.so (no GPU needed to load) and bind every symbol."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise RuntimeError(
                f"gato_amd native library not found at {p}; build it with "
                "`python gato_amd/build.py` (hipcc, gfx950). There is no CPU fallback.")
        lib = C.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def lib() -> C.CDLL:
    return _lib if _lib is not None else load()


def check(status: int) -> None:
    if status == GATO_OK:
        return
    msg = lib().gato_last_error().decode(errors="replace")
    if status == GATO_ERR_VALUE:
        raise ValueError(msg)
    if status == GATO_ERR_KEY:
        raise KeyError(msg)
    raise RuntimeError(f"gato_amd error {status}: {msg}")


def require_cuda(t: torch.Tensor, name: str = "tensor") -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU (ROCm) tensor; gato_amd has no CPU path")


def dtype_code(dt: torch.dtype) -> int:
    if dt == torch.float32:
        return F32
    if dt == torch.float64:
        return F64
    raise ValueError(f"unsupported dtype {dt}; expected float32 or float64")


def torch_dtype(code: int) -> torch.dtype:
    return torch.float32 if code == F32 else torch.float64


def device_index(device=None) -> int:
    """Index of ``device`` (the current device when None or index-less)."""
    if device is not None:
        d = torch.device(device)
        if d.index is not None:
            return d.index
    return torch._C._cuda_getDevice()


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_ptr(device=None) -> int:
    """Raw handle of the current stream of a device (the current device by
    default).  torch's own raw-stream query where it has one: building a
    Stream object per call cost about 7 us, and the SCF loop asks hundreds of
    times per solve."""
    idx = device_index(device)
    if _raw_stream is not None:
        return _raw_stream(idx)
    return torch.cuda.current_stream(idx).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


_ws: dict = {}


def workspace(device=None) -> torch.Tensor:
    """Reduction workspace (partials and staging) of the current stream on a
    device, zero-filled once.

    Reductions store their wave or block partials here and a finalize kernel
    on the same stream sums them, so the buffer needs no reset between
    calls.  Calls that overlap on different streams must not share it, so
    each stream gets its own."""
    idx = device_index(device)
    key = (idx, stream_ptr(idx))
    w = _ws.get(key)
    if w is None:
        nbytes = int(lib().gato_reduce_workspace_bytes())
        w = torch.zeros(nbytes, dtype=torch.uint8, device=torch.device("cuda", idx))
        _ws[key] = w
    return w


_retired_lock = threading.Lock()
_retired: list | None = None


def retire(obj) -> None:
    """Drop a plan evicted from a cache.  Inside ``deferred_release()`` it is
    kept alive until the scope ends: a plan's destructor synchronises the
    whole device, which would stall every other stream's work."""
    with _retired_lock:
        if _retired is not None:
            _retired.append(obj)


@contextlib.contextmanager
def deferred_release():
    """Hold every plan evicted inside the scope until its end, then wait for
    the device and release them together."""
    global _retired
    with _retired_lock:
        outer = _retired is not None
        if not outer:
            _retired = []
    try:
        yield
    finally:
        if not outer:
            with _retired_lock:
                held, _retired = _retired, None
            if held:
                torch.cuda.synchronize()
            del held


def double_array(values) -> C.Array:
    vals = [float(v) for v in values]
    return (C.c_double * max(len(vals), 1))(*vals)


def int_array(values) -> C.Array:
    vals = [int(v) for v in values]
    return (C.c_int * max(len(vals), 1))(*vals)
