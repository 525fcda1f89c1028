"""Host-side basis construction and workload generators (no GPU)."""

import math

import numpy as np
import pytest

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set, even_tempered_dict, normalize_shell, primitive_norm
from jax_qc.basis.library import STO_3G
from oracle import mmd


def test_primitive_norm_formulas():
    # test_basis.py:85-99
    for a in (0.1, 1.0, 10.0):
        assert primitive_norm(a, 0) == pytest.approx((2 * a / math.pi) ** 0.75)
    for a in (0.5, 1.0, 2.5):
        assert primitive_norm(a, 1) == pytest.approx((2 * a / math.pi) ** 0.75 * math.sqrt(4 * a))


def test_normalized_shells_have_unit_self_overlap():
    # test_basis.py:102-133, checked with the oracle's overlap integrals
    for l in (0, 1, 2):
        a = np.array([5.0, 1.0, 0.3])
        c = normalize_shell(a, np.array([0.1, 0.4, 0.5]), l)
        S = sum(c[i] * c[j] * mmd.primitive_overlap_block(a[i], np.zeros(3), l, a[j], np.zeros(3), l)
                for i in range(3) for j in range(3))
        assert S[0, 0] == pytest.approx(1.0, abs=1e-12)


def test_sto3g_first_row_follows_zeta_scaling():
    # STO-3G exponents are zeta^2 times universal 1s / 2sp fits: the ratios
    # between elements must be exactly (zeta_X / zeta_Y)^2.
    zeta1s = {"1": 1.24, "6": 5.67, "7": 6.67, "8": 7.66, "9": 8.65}
    zeta2sp = {"6": 1.72, "7": 1.95, "8": 2.25, "9": 2.55}
    h = np.array([float(x) for x in STO_3G["elements"]["1"]["electron_shells"][0]["exponents"]])
    for z, zeta in zeta1s.items():
        e = np.array([float(x) for x in STO_3G["elements"][z]["electron_shells"][0]["exponents"]])
        np.testing.assert_allclose(e / h, (zeta / 1.24) ** 2, rtol=2e-9)
    c = np.array([float(x) for x in STO_3G["elements"]["6"]["electron_shells"][1]["exponents"]])
    for z, zeta in zeta2sp.items():
        e = np.array([float(x) for x in STO_3G["elements"][z]["electron_shells"][1]["exponents"]])
        np.testing.assert_allclose(e / c, (zeta / 1.72) ** 2, rtol=2e-9)


def test_shell_order_and_sp_split():
    bs = build_basis_set(M.h2o_tier2(), "6-31g**")
    ls = [s.angular_momentum for s in bs.shells if s.atom_index == 0]
    assert ls == [0, 0, 0, 1, 1, 2]
    ex = [float(np.max(s.exponents)) for s in bs.shells if s.atom_index == 0 and s.angular_momentum == 0]
    assert ex == sorted(ex, reverse=True)
    assert bs.n_basis == 24


def test_water_cluster_sizes():
    mol = M.water_cluster((2, 2, 2), seed=0)
    bs = build_basis_set(mol, "even-tempered")
    assert len(mol.elements) == 24 and mol.n_electrons == 80
    assert len(bs.shells) == 128 and bs.n_basis == 240
    mol2 = M.water_cluster((2, 2, 2), seed=0)
    np.testing.assert_array_equal(mol.coords, mol2.coords)
    d = np.linalg.norm(mol.coords[:, None] - mol.coords[None], axis=-1) + np.eye(24) * 99
    assert d.min() > 1.0  # Bohr


def test_even_tempered_is_seed_stable():
    a, b = even_tempered_dict(0), even_tempered_dict(0)
    assert a == b
    assert len(a["elements"]["8"]["electron_shells"]) == 8
    assert len(a["elements"]["1"]["electron_shells"]) == 4


def test_random_molecules_closed_shell():
    mols = M.random_molecules(32, seed=3)
    for m in mols:
        assert m.n_electrons % 2 == 0 and 2 <= len(m.elements) <= 13
        d = np.linalg.norm(m.coords[:, None] - m.coords[None], axis=-1) + np.eye(len(m.elements)) * 99
        assert d.min() >= 0.9 * 1.8897 - 1e-9


def test_make_molecule_validation():
    from jax_qc.core.types import make_molecule

    with pytest.raises(ValueError):
        make_molecule(("H",), [[0, 0, 0]], [1], spin=0)
    m = make_molecule(("O", "O"), [[0, 0, 0], [0, 0, 2.28]], [8, 8], spin=2)
    assert (m.n_alpha, m.n_beta) == (9, 7)
