"""HOMO/LUMO analysis (jax_qc/properties/orbital_analysis.py:18-101): RHF
orbitals doubly occupied up to N/2; UHF uses the alpha orbital energies
held in ``result.state``."""

from __future__ import annotations

from typing import NamedTuple, Optional

import torch


class OrbitalInfo(NamedTuple):
    orbital_energies: torch.Tensor
    occupations: torch.Tensor
    homo_energy: float
    lumo_energy: Optional[float]
    homo_lumo_gap: Optional[float]
    homo_index: int
    lumo_index: Optional[int]


def orbital_analysis(result, mol) -> OrbitalInfo:
    orb_e = result.state.orbital_energies
    n = orb_e.shape[-1]
    n_occ = mol.n_electrons // 2 if mol.spin == 0 else mol.n_alpha
    occ = torch.zeros(n, dtype=torch.float64, device=orb_e.device)
    occ[:n_occ] = 2.0 if mol.spin == 0 else 1.0
    homo = float(orb_e[n_occ - 1])
    if n_occ < n:
        lumo = float(orb_e[n_occ])
        return OrbitalInfo(orb_e, occ, homo, lumo, lumo - homo, n_occ - 1, n_occ)
    return OrbitalInfo(orb_e, occ, homo, None, None, n_occ - 1, None)
