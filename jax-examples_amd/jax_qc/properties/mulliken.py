"""Mulliken populations and charges (jax_qc/properties/mulliken.py:21-90):
P_mu = (D S)_mu,mu summed per atom; q_A = Z_A - P_A."""

from __future__ import annotations

import numpy as np
import torch


def mulliken_populations(D, S, basis_to_atom, n_atoms: int) -> torch.Tensor:
    ps = torch.diagonal(D @ S)
    idx = torch.as_tensor(np.asarray(basis_to_atom, dtype=np.int64), device=ps.device)
    return torch.zeros(n_atoms, dtype=ps.dtype, device=ps.device).index_add_(0, idx, ps)


def mulliken_charges(result, mol, basis) -> torch.Tensor:
    return mulliken_analysis(result, mol, basis)[1]


def mulliken_analysis(result, mol, basis):
    """(populations, charges) per atom."""
    D, S = result.state.density, result.S
    pops = mulliken_populations(D, S, basis.basis_to_atom, len(mol.elements))
    Z = torch.as_tensor(np.asarray(mol.atomic_numbers, dtype=np.float64), device=pops.device)
    return pops, Z - pops
