"""Properties on the device path: exact dipole integrals against the
quadrature restatement (oracle/properties.py), origin independence of a
neutral molecule's dipole, Mulliken charges against the reference's formula
(jax_qc/properties/mulliken.py:21-90) on oracle S and D, orbital analysis."""

import numpy as np
import pytest
import torch

import jax_qc
from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from jax_qc.core.types import CalcConfig
from jax_qc.integrals.one_electron import compute_dipole_integrals
from jax_qc.properties import (
    dipole_moment,
    dipole_moment_debye,
    mulliken_analysis,
    orbital_analysis,
)
from oracle import mmd, scf
from oracle.properties import moment_matrices

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g"])
def test_dipole_integrals_match_quadrature(basis):
    mol = M.water_monomer()
    bs = build_basis_set(mol, basis)
    for O in ((0.0, 0.0, 0.0), (0.4, -1.2, 2.5)):
        print(f"[case] dipole ints {basis} origin {O}", flush=True)
        got = compute_dipole_integrals(bs, O).cpu().numpy()
        ref = moment_matrices(bs, O)[1:]
        np.testing.assert_allclose(got, ref, atol=1e-12)


def test_dipole_origin_independent_and_mulliken():
    mol = M.water_monomer()
    bs = build_basis_set(mol, "sto-3g")
    ints = jax_qc.compute_integrals(mol, bs)
    res = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf", scf_conv=1e-10))
    mu0 = dipole_moment(res, mol, bs).cpu().numpy()
    O = np.array([1.0, -0.5, 0.7])
    D = res.state.density
    MO = compute_dipole_integrals(bs, O).to(D.device)
    Z = np.asarray(mol.atomic_numbers, float)
    muO = Z @ (np.asarray(mol.coords) - O) - torch.einsum("mn,kmn->k", D, MO).cpu().numpy()
    np.testing.assert_allclose(muO, mu0, atol=1e-10)
    print("H2O STO-3G dipole (Debye):", dipole_moment_debye(res, mol, bs), flush=True)
    assert abs(mu0[0]) < 1e-10 and abs(mu0[1]) < 1e-10  # the molecule lies in the yz plane, C2 axis z
    pops, q = mulliken_analysis(res, mol, bs)
    ref = scf.run_rhf(mol, mmd.compute_integrals(mol, bs), conv=1e-10)
    ps = np.diag(np.asarray(ref["D"]) @ mmd.one_electron(bs, "S"))
    pops_ref = np.zeros(3)
    np.add.at(pops_ref, np.asarray(bs.basis_to_atom), ps)
    np.testing.assert_allclose(pops.cpu().numpy(), pops_ref, atol=1e-8)
    assert abs(float(q.sum()) - mol.charge) < 1e-10
    assert q[0] < 0 < q[1]
    info = orbital_analysis(res, mol)
    assert info.homo_index == 4 and info.lumo_index == 5 and info.homo_lumo_gap > 0
