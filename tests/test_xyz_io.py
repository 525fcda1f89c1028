"""XYZ read/write (jax_qc/io/xyz.py:32-119): the cases of the reference's
tests/test_xyz.py (parse, units, charge/spin, malformed input, round trip)
plus a round trip of a generated water cluster."""

import numpy as np
import pytest

from jax_qc import molecules as M
from jax_qc.core.constants import ANGSTROM_TO_BOHR, BOHR_TO_ANGSTROM
from jax_qc.core.types import make_molecule
from jax_qc.io import parse_xyz_string, read_xyz, write_xyz
from jax_qc.io.xyz import format_xyz

WATER = "3\nwater, Angstrom\nO   0.000000   0.000000   0.117176\n" \
        "H   0.000000   0.757087  -0.468706\nH   0.000000  -0.757087  -0.468706\n"


def test_parse_water_angstrom_to_bohr():
    mol = parse_xyz_string(WATER)
    assert mol.elements == ("O", "H", "H")
    assert list(np.asarray(mol.atomic_numbers)) == [8, 1, 1]
    assert mol.n_electrons == 10
    assert float(mol.coords[0, 2]) == pytest.approx(0.117176 * ANGSTROM_TO_BOHR, abs=1e-12)


def test_parse_bohr_lowercase_symbols_and_blank_lines():
    mol = parse_xyz_string("2\n\n\nh 0.0 0.0 0.0\n\nh 0 0 1.0\n", unit="bohr")
    assert mol.elements == ("H", "H")
    assert float(mol.coords[1, 2]) == 1.0


def test_parse_charge_and_spin():
    mol = parse_xyz_string("2\n\nO 0 0 0\nO 0 0 2.282\n", unit="bohr", spin=2)
    assert (mol.spin, mol.n_alpha, mol.n_beta) == (2, 9, 7)


@pytest.mark.parametrize("text", ["", "x\ncomment\nH 0 0 0\n", "3\n\nH 0 0 0\nH 0 0 1\n",
                                  "1\n\nH 0 0\n"],
                         ids=["empty", "bad-count", "truncated", "malformed"])
def test_parse_rejects_bad_input(text):
    with pytest.raises(ValueError):
        parse_xyz_string(text, unit="bohr")


def test_unknown_unit_rejected(tmp_path):
    with pytest.raises(ValueError):
        parse_xyz_string(WATER, unit="nm")
    with pytest.raises(ValueError):
        write_xyz(tmp_path / "x.xyz", parse_xyz_string(WATER), unit="pm")


def test_round_trip_file(tmp_path):
    p = tmp_path / "w.xyz"
    p.write_text(WATER, encoding="utf-8")
    mol = read_xyz(p)
    out = tmp_path / "w2.xyz"
    write_xyz(out, mol, comment="round trip")
    mol2 = read_xyz(out)
    assert mol2.elements == mol.elements
    np.testing.assert_allclose(np.asarray(mol2.coords), np.asarray(mol.coords), atol=1e-9)
    assert out.read_text().splitlines()[1] == "round trip"


def test_write_units():
    h2 = make_molecule(("H", "H"), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]]),
                       np.array([1, 1], dtype=np.int32))
    bohr = format_xyz(h2, unit="bohr").splitlines()
    ang = format_xyz(h2).splitlines()
    assert bohr[0] == "2"
    assert float(bohr[3].split()[-1]) == pytest.approx(1.4)
    assert float(ang[3].split()[-1]) == pytest.approx(1.4 * BOHR_TO_ANGSTROM, abs=1e-9)


def test_round_trip_water_cluster_with_spin(tmp_path):
    mol = M.water_cluster((2, 2, 1), seed=1, spin=2)
    p = tmp_path / "c.xyz"
    write_xyz(p, mol, unit="bohr")
    back = read_xyz(p, unit="bohr", spin=2)
    assert back.elements == tuple(mol.elements)
    np.testing.assert_allclose(np.asarray(back.coords), np.asarray(mol.coords), atol=1e-10)
    assert (back.n_alpha, back.n_beta) == (mol.n_alpha, mol.n_beta)
