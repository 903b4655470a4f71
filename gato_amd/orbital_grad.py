"""Orbital gradients behind torch.autograd, for direct minimisation.

The reference differentiates only the nuclear positions (jax.grad of the BO
energy at fixed orbitals, water.py:221-224).  These adjoints differentiate
with respect to the orbitals themselves:

- ``rayleigh_quotient(H, psi)``: R = <psi|H|psi> / <psi|psi>
  (hamiltonian.py:33-35), with dR/dpsi = 2 dV (H psi - R psi) / <psi|psi>;
- ``rhf_energy_ad(stack, V_ext, grid, ...)``: the closed-shell RHF energy
  (scf.py:234-286) of an (n, N, N, N) orbital stack, with
  dE/dphi_i = 4 dV F[phi] phi_i, where F = h + V_H[rho] - K[phi] is the Fock
  operator of the same orbitals (scf.py:114-158).  For one orbital this is
  4 (h + J) phi: E = 2h + J, where J = (phi phi|phi phi) and K phi = J phi.

Both backward passes recompute the operator with the native kernels: one
fused stencil, or the Fock apply with its exchange FFT solves.  Nothing from
the forward pass is stored apart from the input fields.
``rayleigh_descent`` minimises R by normalised steepest-descent steps taken
through the adjoint.
"""
from __future__ import annotations

from typing import Callable

import torch

from . import _native as nat
from .grid import Grid3D, inner_product, norm_sq
from .scf import RHFFock, decompose_rhf_energy, to_grid


class _RayleighQuotient(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi, H):
        psi = nat.contig(psi, "psi")
        den = norm_sq(psi, H.grid)
        R = inner_product(psi, H.apply(psi), H.grid) / den
        ctx.save_for_backward(psi)
        ctx.H, ctx.R, ctx.den = H, float(R), float(den)
        return R

    @staticmethod
    def backward(ctx, g):
        (psi,) = ctx.saved_tensors
        H = ctx.H
        c = 2.0 * H.grid.dV * float(g) / ctx.den
        grad = H.apply(psi)                              # recomputed, not stored
        nat.axpby(-c * ctx.R, psi, c, grad)              # c (H psi - R psi)
        return grad, None


def rayleigh_quotient(H, psi: torch.Tensor) -> torch.Tensor:
    """<psi|H|psi> / <psi|psi> for a ``Hamiltonian`` H, differentiable in psi."""
    return _RayleighQuotient.apply(psi, H)


class _RHFEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, stack, V_ext, grid, boundary, order, epsilon, V_nl_apply):
        stack = nat.contig(stack, "orbitals")
        t = decompose_rhf_energy(to_grid(stack), V_ext, grid, boundary=boundary, order=order,
                                 epsilon=epsilon, V_nl_apply=V_nl_apply)
        ctx.save_for_backward(stack)
        ctx.args = (V_ext, grid, boundary, order, epsilon, V_nl_apply)
        return torch.tensor(t.total, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        (stack,) = ctx.saved_tensors
        V_ext, grid, boundary, order, epsilon, V_nl_apply = ctx.args
        F = RHFFock(grid=grid, V_ext=V_ext, orbitals=to_grid(stack), boundary=boundary,
                    order=order, epsilon=epsilon, V_nl_apply=V_nl_apply)
        grad = torch.empty_like(stack)
        scale = 4.0 * grid.dV * float(g)
        for i in range(stack.shape[0]):
            grad[i].copy_(F.apply(stack[i]))             # F phi_i, recomputed
        nat.axpby(0.0, grad, scale, grad)
        return grad, None, None, None, None, None, None


def rhf_energy_ad(stack: torch.Tensor, V_ext: torch.Tensor, grid: Grid3D, *,
                  boundary: str = "dirichlet", order: int = 4, epsilon: float | None = None,
                  V_nl_apply: Callable | None = None) -> torch.Tensor:
    """The closed-shell RHF electronic energy of the (n, N, N, N) orbital stack
    (a float64 scalar on the host), differentiable in the orbitals."""
    return _RHFEnergy.apply(stack, V_ext, grid, boundary, order, epsilon, V_nl_apply)


def rayleigh_descent(H, psi0: torch.Tensor, n_steps: int = 500, tau: float | None = None,
                     log_every: int = 50):
    """Lowest eigenpair of H by normalised steepest descent on the Rayleigh
    quotient, with the gradient from ``rayleigh_quotient``'s adjoint:
    psi <- normalise(psi - tau (H psi - R psi)).  tau defaults to h^2 / 8,
    inside the stability bound 2 / (lambda_max - lambda_min) for both stencil
    orders.  Returns (psi, [(step, R), ...])."""
    grid = H.grid
    if tau is None:
        tau = grid.h ** 2 / 8.0
    psi = nat.normalize(nat.contig(psi0, "psi0"), grid.dV)
    hist = []
    for k in range(n_steps):
        x = psi.detach().requires_grad_(True)
        R = rayleigh_quotient(H, x)
        (g,) = torch.autograd.grad(R, x)
        if k % log_every == 0 or k == n_steps - 1:
            hist.append((k, float(R.detach())))
        # for a normalised psi the gradient is 2 dV (H psi - R psi)
        nat.axpby(-tau / (2.0 * grid.dV), g, 1.0, psi)
        psi = nat.normalize(psi, grid.dV)
    return psi, hist
