"""Overlap, kinetic and nuclear-attraction matrices on the GPU
(replace jax_qc/integrals/overlap.py:67-85, kinetic.py:42-59,
nuclear.py:62-81) and the nuclear repulsion energy (nuclear.py:84-98)."""

from __future__ import annotations

import numpy as np
import torch

from jax_qc import _native as nat
from jax_qc.integrals import engine as eng


def _one(basis, which: int, mol=None, pair_table=None, origin=None) -> torch.Tensor:
    dev = nat.device()
    nat.ensure_init(dev)
    pt = pair_table or eng.PairTable(basis, dev, fp32=False)
    n = pt.nbf
    M = torch.zeros((n, n), dtype=torch.float64, device=dev)
    if origin is not None:  # dipole integrals: the origin goes where the nuclei would
        xyz = torch.as_tensor(np.asarray(origin, dtype=np.float64).reshape(3), device=dev)
        zz = None
        nat_atoms = 0
    elif mol is not None:
        xyz = torch.as_tensor(np.ascontiguousarray(mol.coords, dtype=np.float64), device=dev)
        zz = torch.as_tensor(np.asarray(mol.atomic_numbers, dtype=np.float64), device=dev)
        nat_atoms = len(zz)
    else:
        xyz = zz = None
        nat_atoms = 0
    for c in range(6):
        lst = pt.class_pairs[c]
        if len(lst) == 0:
            continue
        ld = torch.from_numpy(lst).to(dev)
        nat.call("qc_one_electron", which, c, nat.ptr(pt.pairs64), nat.ptr(pt.sp_prim_off),
                 nat.ptr(pt.sp_nprim), nat.ptr(pt.sp_ao), nat.ptr(ld), len(lst), nat_atoms,
                 nat.ptr(xyz), nat.ptr(zz), n, nat.ptr(M), nat.stream_ptr())
    return M


def compute_overlap_matrix(basis, pair_table=None) -> torch.Tensor:
    return _one(basis, 0, pair_table=pair_table)


def compute_kinetic_matrix(basis, pair_table=None) -> torch.Tensor:
    return _one(basis, 1, pair_table=pair_table)


def compute_nuclear_matrix(basis, mol, pair_table=None) -> torch.Tensor:
    return _one(basis, 2, mol, pair_table=pair_table)


def nuclear_repulsion_energy(mol) -> float:
    """E_nuc = sum_{A<B} Z_A Z_B / R_AB; raises on coincident atoms."""
    c = np.asarray(mol.coords, dtype=np.float64)
    Z = np.asarray(mol.atomic_numbers, dtype=np.float64)
    d = np.linalg.norm(c[:, None, :] - c[None, :, :], axis=-1)
    iu = np.triu_indices(len(Z), 1)
    if np.any(d[iu] == 0.0):
        i, j = [int(x[np.argmin(d[iu])]) for x in iu]
        raise ValueError(f"Atoms {i} and {j} are coincident; cannot compute E_nuc.")
    return float(np.sum(np.outer(Z, Z)[iu] / d[iu]))


def compute_dipole_integrals(basis, origin=(0.0, 0.0, 0.0), pair_table=None) -> torch.Tensor:
    """(3, n, n) exact dipole integrals <mu| r_k - O_k |nu> (the reference
    approximates them, jax_qc/properties/dipole.py:26-89)."""
    return torch.stack([_one(basis, 3 + k, pair_table=pair_table, origin=origin)
                        for k in range(3)])
