"""The quadrature restatement of moment integrals (oracle/properties.py),
pinned on the CPU: its overlap equals the C oracle's overlap (the same
coefficient convention), a single s-s primitive pair has the closed form
<s_A| x - O |s_B> = (P_x - O_x) S, and moving the origin subtracts O S."""

import numpy as np

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from jax_qc.core.types import make_molecule
from oracle import mmd
from oracle.properties import _moment_1d, moment_matrices


def test_quadrature_overlap_equals_oracle_overlap():
    mol = M.water_monomer()
    bs = build_basis_set(mol, "sto-3g")
    S = moment_matrices(bs)[0]
    np.testing.assert_allclose(S, mmd.one_electron(bs, "S"), atol=1e-13)


def test_ss_first_moment_closed_form():
    a, b, A, B, O = 0.7, 1.3, 0.2, -0.9, 0.35
    p = a + b
    P = (a * A + b * B) / p
    S = np.exp(-a * b / p * (A - B) ** 2) * np.sqrt(np.pi / p)
    assert abs(_moment_1d(a, A, 0, b, B, 0, O, 0) - S) < 1e-15
    assert abs(_moment_1d(a, A, 0, b, B, 0, O, 1) - (P - O) * S) < 1e-15


def test_origin_shift_subtracts_origin_times_overlap():
    mol = make_molecule(("H", "F"), np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 1.7]]), [1, 9])
    bs = build_basis_set(mol, "sto-3g")
    m0 = moment_matrices(bs)
    O = np.array([0.5, -1.0, 2.0])
    m1 = moment_matrices(bs, O)
    for k in range(3):
        np.testing.assert_allclose(m1[1 + k], m0[1 + k] - O[k] * m0[0], atol=1e-12)
