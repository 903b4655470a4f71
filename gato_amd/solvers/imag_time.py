"""Imaginary-time propagation (reference: solvers/imag_time.py:24-65).

First-order Euler step psi <- psi - dt H psi, then renormalise; dt defaults
to h^2/10.  The steps between two logged energies run in the library
(gato_imag_time_steps) as one stencil launch per step,
    psi' = psi + (dt/2) lap(psi) - dt V psi,
which also writes the block partials of <psi', psi'>; the next launch sums
them and scales its own result by the inverse norm, so each step is
normalised without a pass of its own, and one final pass normalises the
last.  Two device buffers ping-pong, so a step allocates nothing -- the
native form of the reference's donated update.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from .. import _lib as L
from .. import _native as nat


@dataclass
class ImagTimeHistory:  # solvers/imag_time.py:24-27
    steps: list
    energies: list


def imaginary_time(H, psi0: torch.Tensor, dt: float | None = None, n_steps: int = 2000,
                   log_every: int = 50):
    """Propagate psi0 in imaginary time towards the ground state of H
    (imag_time.py:30-65).  H is a ``Hamiltonian`` (grid, V, boundary, order).
    Returns (psi, ImagTimeHistory) with the Rayleigh quotient recorded every
    ``log_every`` steps and at the last step."""
    grid = H.grid
    if dt is None:
        dt = grid.h ** 2 / 10.0
    psi0 = nat.contig(psi0, "psi0")
    dtype = psi0.dtype
    n = psi0.numel()
    lib = L.lib()
    code = L.dtype_code(dtype)
    ws = L.workspace(psi0.device).data_ptr()
    sp = L.stream_ptr()
    Vs = H._V_as(dtype).clone()
    nat.axpby(0.0, Vs, -float(dt), Vs)  # -dt V
    a = torch.empty_like(psi0)
    b = torch.empty_like(psi0)

    def normalise_into(src: torch.Tensor, dst: torch.Tensor) -> None:
        # dst = src / sqrt(dV <src, src>): norm partials, then one fused
        # finalize-and-scale launch
        L.check(lib.gato_normalize(code, n, src.data_ptr(), grid.dV, None, dst.data_ptr(), ws, sp))

    normalise_into(psi0, a)  # imag_time.py:55
    steps_hist: list = []
    energies: list = []
    order, bnd = nat.check_order(H.order), nat.boundary_code(H.boundary)
    # the energy is logged after step i when i % log_every == 0 and after the
    # last step (imag_time.py:57-60); the steps up to each such point run as
    # one call
    marks = [i for i in range(n_steps) if i % log_every == 0 or i == n_steps - 1]
    done = 0
    for i in marks:
        k = i + 1 - done
        L.check(lib.gato_imag_time_steps(code, grid.N, grid.h, order, bnd, float(dt),
                                         Vs.data_ptr(), a.data_ptr(), b.data_ptr(), k, grid.dV,
                                         None, ws, sp))
        if k % 2:
            a, b = b, a
        done = i + 1
        steps_hist.append(i)
        energies.append(float(H.rayleigh(a)))
    return a, ImagTimeHistory(steps=steps_hist, energies=energies)
