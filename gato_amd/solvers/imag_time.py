"""Imaginary-time propagation (reference: solvers/imag_time.py:24-65).

First-order Euler step psi <- psi - dt H psi, then renormalise; dt defaults
to h^2/10.  One step is one fused stencil launch,
    psi' = psi + (dt/2) lap(psi) - dt V psi
(the +psi term enters through the stencil's subtract operand with
coefficient -1) that also writes the block partials of <psi', psi'>, then
one launch that sums them and scales.  Two device buffers ping-pong, so a step allocates nothing -- the
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
    minus_one = torch.tensor([-1.0], dtype=torch.float64, device=psi0.device)
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
    for i in range(n_steps):
        # b = a + (dt/2) lap(a) - dt V a, renormalised: the stencil writes the
        # norm's block partials, one more launch sums them and scales
        L.check(lib.gato_stencil_normalize(code, grid.N, grid.h, order, bnd, 0.5 * dt,
                                           a.data_ptr(), Vs.data_ptr(), a.data_ptr(),
                                           minus_one.data_ptr(), b.data_ptr(), grid.dV, None,
                                           ws, sp))
        a, b = b, a
        if (i % log_every == 0) or (i == n_steps - 1):
            steps_hist.append(i)
            energies.append(float(H.rayleigh(a)))
    return a, ImagTimeHistory(steps=steps_hist, energies=energies)
