"""Nuclear geometry (reference: geometry.py:1-91).

Positions (K, 3) bohr and charges (K,) are small float64 host tensors; the
position tensor may carry ``requires_grad`` so that Hellmann-Feynman forces
come out of ``torch.autograd`` (the adjoint of the field terms runs on the
GPU, see ``gato_amd.forces``).
"""
from __future__ import annotations

import math
from typing import NamedTuple

import torch


def _t(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(torch.float64)
    return torch.as_tensor(x, dtype=torch.float64)


class Nuclei(NamedTuple):
    """K nuclei (geometry.py:20-28)."""

    positions: torch.Tensor  # (K, 3)
    charges: torch.Tensor    # (K,)

    @property
    def n(self) -> int:
        return int(self.charges.shape[0])


def make_nuclei(positions, charges) -> Nuclei:
    return Nuclei(_t(positions).reshape(-1, 3), _t(charges).reshape(-1))


def diatomic_along_z(R: float, Z1: float, Z2: float | None = None) -> Nuclei:
    """Two nuclei at (0, 0, -R/2) and (0, 0, +R/2) (geometry.py:31-45)."""
    if Z2 is None:
        Z2 = Z1
    return make_nuclei([[0.0, 0.0, -R / 2], [0.0, 0.0, R / 2]], [Z1, Z2])


def bond_length(nuclei: Nuclei, i: int, j: int) -> torch.Tensor:  # geometry.py:48-50
    return torch.linalg.norm(nuclei.positions[i] - nuclei.positions[j])


def bond_angle(nuclei: Nuclei, i: int, j: int, k: int) -> torch.Tensor:  # geometry.py:53-58
    a = nuclei.positions[i] - nuclei.positions[j]
    b = nuclei.positions[k] - nuclei.positions[j]
    c = torch.dot(a, b) / (torch.linalg.norm(a) * torch.linalg.norm(b))
    return torch.arccos(torch.clamp(c, -1.0, 1.0))


def nuclear_repulsion(nuclei: Nuclei) -> torch.Tensor:
    """E_nn = sum_{i<j} Z_i Z_j / R_ij (geometry.py:61-75); differentiable."""
    K = nuclei.n
    P = nuclei.positions
    if K < 2:
        return torch.zeros((), dtype=torch.float64)
    i, j = torch.triu_indices(K, K, offset=1)
    r = torch.linalg.norm(P[i] - P[j], dim=-1)
    return torch.sum(nuclei.charges[i] * nuclei.charges[j] / r)


def center_of_mass(nuclei: Nuclei) -> torch.Tensor:  # geometry.py:78-83
    Z = nuclei.charges
    return (Z[:, None] * nuclei.positions).sum(dim=0) / Z.sum()


def recenter(nuclei: Nuclei) -> Nuclei:  # geometry.py:86-91
    return Nuclei(nuclei.positions - center_of_mass(nuclei), nuclei.charges)
