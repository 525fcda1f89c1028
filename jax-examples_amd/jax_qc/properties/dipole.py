"""Dipole moment (jax_qc/properties/dipole.py:92-150): mu = sum_A Z_A R_A -
Tr(D M_k), origin at (0, 0, 0) as in the reference.  M_k are the exact
dipole integrals from the device kernel (qc_one_electron, which = 3..5);
the reference uses an overlap-times-weighted-centre approximation
(dipole.py:26-89), which is why results differ from it for p and d
functions."""

from __future__ import annotations

import numpy as np
import torch

from jax_qc.integrals.one_electron import compute_dipole_integrals

AU_TO_DEBYE: float = 2.541746473


def _dipole_integrals(basis, S=None, pair_table=None) -> torch.Tensor:
    """(3, n, n) <mu| r_k |nu>; S is accepted for the reference's signature
    and unused (the integrals are computed, not approximated from S)."""
    return compute_dipole_integrals(basis, (0.0, 0.0, 0.0), pair_table)


def nuclear_dipole(mol) -> torch.Tensor:
    Z = np.asarray(mol.atomic_numbers, dtype=np.float64)
    return torch.from_numpy(Z @ np.asarray(mol.coords, dtype=np.float64))


def electronic_dipole(D: torch.Tensor, dipole_ints: torch.Tensor) -> torch.Tensor:
    return -torch.einsum("mn,kmn->k", D, dipole_ints)


def dipole_moment(result, mol, basis) -> torch.Tensor:
    """(3,) total dipole in atomic units (e Bohr)."""
    D = result.state.density
    M = _dipole_integrals(basis).to(D.device)
    return nuclear_dipole(mol).to(D.device) + electronic_dipole(D, M)


def dipole_moment_debye(result, mol, basis) -> float:
    return float(torch.linalg.norm(dipole_moment(result, mol, basis))) * AU_TO_DEBYE
