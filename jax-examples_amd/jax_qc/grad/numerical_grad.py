"""Full-SCF central finite-difference gradient (replaces
jax_qc/grad/numerical_grad.py:51-94): 6 N_atoms SCF energies at R +- h e."""

from __future__ import annotations

import numpy as np
import torch

from jax_qc.basis import build_basis_set
from jax_qc.core.types import CalcConfig, make_molecule
from jax_qc.integrals import compute_integrals
from jax_qc.scf import run_scf


def displaced(mol, coords):
    return make_molecule(mol.elements, coords, mol.atomic_numbers, mol.charge, mol.spin)


def scf_energy_at(mol, coords, basis_name, config, eri="store", tau=0.0, tau32=None):
    m = displaced(mol, coords)
    bs = build_basis_set(m, basis_name)
    ints = compute_integrals(m, bs, eri=eri, tau=tau, tau32=tau32)
    r = run_scf(m, ints, config)
    return float(r.energy)


def numerical_gradient(mol, basis_name="sto-3g", method="rhf", step_size=1e-4, scf_conv=1e-10,
                       eri="store", tau=0.0, tau32=None, atoms=None):
    """dE/dR by central differences of full SCF energies.  ``atoms``
    restricts the displaced atoms (others are left zero)."""
    config = CalcConfig(method=method, basis=basis_name, scf_conv=scf_conv, max_scf_iter=200)
    c = np.asarray(mol.coords, dtype=np.float64)
    g = np.zeros_like(c)
    for a in (range(len(c)) if atoms is None else atoms):
        for k in range(3):
            cp, cm = c.copy(), c.copy()
            cp[a, k] += step_size
            cm[a, k] -= step_size
            ep = scf_energy_at(mol, cp, basis_name, config, eri, tau, tau32)
            em = scf_energy_at(mol, cm, basis_name, config, eri, tau, tau32)
            g[a, k] = (ep - em) / (2 * step_size)
    return torch.from_numpy(g)
