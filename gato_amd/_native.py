"""Tensor-level wrappers over the C ABI: shape/dtype checks, device pointers,
the current stream, and the shared reduction workspace.  Everything here
launches HIP kernels of libgato_amd.so; nothing computes on the host."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L

BOUNDARIES = {"dirichlet": L.DIRICHLET, "periodic": L.PERIODIC}


def boundary_code(boundary: str) -> int:
    if boundary not in BOUNDARIES:
        raise ValueError(f"Unknown boundary {boundary!r}; expected 'dirichlet' or 'periodic'.")
    return BOUNDARIES[boundary]


def check_order(order: int) -> int:
    if order not in (2, 4):
        raise ValueError(f"Unknown order {order!r}; expected 2 or 4.")
    return int(order)


def field_grid(psi: torch.Tensor) -> tuple[int, int]:
    """(N, batch) of a (..., N, N, N) field stack."""
    if psi.dim() < 3 or not (psi.shape[-1] == psi.shape[-2] == psi.shape[-3]):
        raise ValueError(f"expected (..., N, N, N) fields, got shape {tuple(psi.shape)}")
    N = int(psi.shape[-1])
    batch = int(psi.numel() // (N ** 3)) if psi.numel() else 0
    return N, batch


def contig(t: torch.Tensor, name: str = "field") -> torch.Tensor:
    L.require_cuda(t, name)
    L.dtype_code(t.dtype)
    return t if t.is_contiguous() else t.contiguous()


def stencil(psi: torch.Tensor, h: float, order: int, boundary: str, c_lap: float,
            V: torch.Tensor | None = None, out: torch.Tensor | None = None,
            accumulate: bool = False, sub: torch.Tensor | None = None,
            sub_coef: torch.Tensor | None = None) -> torch.Tensor:
    psi = contig(psi, "psi")
    N, batch = field_grid(psi)
    order = check_order(order)
    b = boundary_code(boundary)
    if out is None:
        out = torch.empty_like(psi)
    if V is not None:
        V = contig(V, "V")
        if V.dtype != psi.dtype or V.shape != psi.shape[-3:]:
            raise ValueError("V must be one (N, N, N) field of the same dtype as psi")
    L.check(L.lib().gato_stencil_apply(
        L.dtype_code(psi.dtype), N, batch, float(h), order, b, float(c_lap), psi.data_ptr(),
        L.ptr(V), int(accumulate), L.ptr(sub), L.ptr(sub_coef), out.data_ptr(), L.stream_ptr()))
    return out


def dot_batched(A: torch.Tensor, B: torch.Tensor, nk: int, lda: int, ldb: int,
                C_: torch.Tensor | None = None, scale: float = 1.0,
                out: torch.Tensor | None = None, n: int | None = None) -> torch.Tensor:
    """out[k] = scale * sum_i A[k*lda+i] B[k*ldb+i] C[i] (float64 device vector)."""
    A = contig(A, "A")
    B = contig(B, "B")
    if A.dtype != B.dtype or (C_ is not None and C_.dtype != A.dtype):
        raise ValueError("dot operands must share one dtype")
    if n is None:
        n = B.numel() if ldb == 0 else ldb
    if out is None:
        out = torch.empty(nk, dtype=torch.float64, device=A.device)
    ws = L.workspace(A.device)
    L.check(L.lib().gato_dot_batched(
        L.dtype_code(A.dtype), int(n), int(nk), A.data_ptr(), int(lda), B.data_ptr(), int(ldb),
        L.ptr(C_), None, None, float(scale), out.data_ptr(), ws.data_ptr(), L.stream_ptr()))
    return out


def density_mix(orbitals_knnn: torch.Tensor, mixing: float, rho_old: torch.Tensor | None,
                out: torch.Tensor | None = None) -> torch.Tensor:
    orb = contig(orbitals_knnn, "orbitals")
    N = int(orb.shape[-1])
    n = N ** 3
    k = int(orb.numel() // n)
    if out is None:
        out = torch.empty(orb.shape[-3:], dtype=orb.dtype, device=orb.device)
    L.check(L.lib().gato_density_mix(L.dtype_code(orb.dtype), n, k, orb.data_ptr(), n,
                                     float(mixing), L.ptr(rho_old), out.data_ptr(),
                                     L.stream_ptr()))
    return out


def axpby(a: float, x: torch.Tensor, b: float, y: torch.Tensor) -> torch.Tensor:
    """y = a x + b y in place."""
    L.check(L.lib().gato_axpby(L.dtype_code(y.dtype), y.numel(), float(a), x.data_ptr(), float(b),
                               y.data_ptr(), L.stream_ptr()))
    return y


def normalize(x: torch.Tensor, scale: float, out: torch.Tensor | None = None,
              norm: torch.Tensor | None = None) -> torch.Tensor:
    """out = x / sqrt(scale <x, x>) (zero when the norm is zero); ``out`` may be
    ``x``; ``norm`` (a 1-element float64 device tensor) receives the norm."""
    x = contig(x, "x")
    if out is None:
        out = torch.empty_like(x)
    L.check(L.lib().gato_normalize(L.dtype_code(x.dtype), x.numel(), x.data_ptr(), float(scale),
                                   None if norm is None else norm.data_ptr(), out.data_ptr(),
                                   L.workspace(x.device).data_ptr(), L.stream_ptr()))
    return out


def convert(x: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    x = contig(x, "x")
    y = torch.empty(x.shape, dtype=dtype, device=x.device)
    L.check(L.lib().gato_convert(L.dtype_code(x.dtype), x.data_ptr(), L.dtype_code(dtype),
                                 y.data_ptr(), x.numel(), L.stream_ptr()))
    return y


def local_potential(N: int, L_box: float, centers: list, dtype: torch.dtype, device=None,
                    out: torch.Tensor | None = None, accumulate: bool = False) -> torch.Tensor:
    """centers: list of 10-tuples [kind, Rx, Ry, Rz, a0..a5]."""
    if out is None:
        dev = device if device is not None else torch.device("cuda")
        out = torch.empty((N, N, N), dtype=dtype, device=dev)
    flat = [float(v) for c in centers for v in c]
    arr = L.double_array(flat)
    L.check(L.lib().gato_local_potential(L.dtype_code(out.dtype), int(N), float(L_box),
                                         len(centers), arr, int(accumulate), out.data_ptr(),
                                         L.stream_ptr()))
    return out


def local_potential_grad(N: int, L_box: float, centers: list, rho: torch.Tensor) -> torch.Tensor:
    rho = contig(rho, "rho")
    g = torch.empty(3 * len(centers), dtype=torch.float64, device=rho.device)
    flat = [float(v) for c in centers for v in c]
    ws = L.workspace(rho.device)
    L.check(L.lib().gato_local_potential_grad(
        L.dtype_code(rho.dtype), int(N), float(L_box), len(centers), L.double_array(flat),
        rho.data_ptr(), g.data_ptr(), ws.data_ptr(), L.stream_ptr()))
    return g.view(len(centers), 3)
