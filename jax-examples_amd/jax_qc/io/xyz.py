"""XYZ files (replaces jax_qc/io/xyz.py:32-119).

Layout: a line holding the atom count, a free comment line, then one
``symbol x y z`` line per atom.  Coordinates are Angstrom by convention and
are held in Bohr inside a Molecule; ``unit='bohr'`` reads and writes them
unconverted.  Same behaviour as the reference: element symbols are
title-cased, blank lines after the comment are ignored, and a missing count,
a short body, a line with fewer than four fields or an unknown unit raise
ValueError.
"""

from __future__ import annotations

import pathlib
from typing import Union

import numpy as np

from jax_qc.core.constants import ANGSTROM_TO_BOHR, BOHR_TO_ANGSTROM, symbol_to_z
from jax_qc.core.types import Molecule, make_molecule

PathLike = Union[str, pathlib.Path]


def _to_bohr_factor(unit: str) -> float:
    u = unit.lower()
    if u == "angstrom":
        return ANGSTROM_TO_BOHR
    if u == "bohr":
        return 1.0
    raise ValueError(f"Unknown unit {unit!r}; expected 'angstrom' or 'bohr'.")


def parse_xyz_string(text: str, unit: str = "angstrom", charge: int = 0,
                     spin: int = 0) -> Molecule:
    """Molecule (Bohr) from the text of an XYZ file."""
    lines = text.splitlines()
    if len(lines) < 2:
        raise ValueError("XYZ input must contain at least a count line and a comment line.")
    head = lines[0].split()
    try:
        natom = int(head[0])
    except (IndexError, ValueError) as exc:
        raise ValueError(f"Invalid atom count on first line: {lines[0]!r}") from exc
    atoms = [ln for ln in lines[2:] if ln.strip()]
    if len(atoms) < natom:
        raise ValueError(
            f"XYZ declares {natom} atoms but only {len(atoms)} coordinate lines were found.")
    symbols, xyz = [], np.empty((natom, 3), dtype=np.float64)
    for i, ln in enumerate(atoms[:natom]):
        fields = ln.split()
        if len(fields) < 4:
            raise ValueError(f"Malformed XYZ coordinate line: {ln!r}")
        symbols.append(fields[0].strip().title())
        xyz[i] = [float(v) for v in fields[1:4]]
    xyz *= _to_bohr_factor(unit)
    Z = np.array([symbol_to_z(s) for s in symbols], dtype=np.int32)
    return make_molecule(tuple(symbols), xyz, Z, charge=charge, spin=spin)


def read_xyz(path: PathLike, unit: str = "angstrom", charge: int = 0, spin: int = 0) -> Molecule:
    """Molecule (Bohr) from an XYZ file."""
    return parse_xyz_string(pathlib.Path(path).read_text(encoding="utf-8"), unit=unit,
                            charge=charge, spin=spin)


def format_xyz(molecule: Molecule, unit: str = "angstrom", comment: str = "") -> str:
    """Text of an XYZ file for ``molecule`` (10 decimals)."""
    xyz = np.asarray(molecule.coords, dtype=np.float64).reshape(-1, 3) / _to_bohr_factor(unit)
    rows = [str(len(molecule.elements)), comment]
    rows += [f"{s:<3s} {x: .10f} {y: .10f} {z: .10f}" for s, (x, y, z) in
             zip(molecule.elements, xyz)]
    return "\n".join(rows) + "\n"


def write_xyz(path: PathLike, molecule: Molecule, unit: str = "angstrom",
              comment: str = "") -> None:
    """Write ``molecule`` as an XYZ file (Angstrom unless unit='bohr')."""
    pathlib.Path(path).write_text(format_xyz(molecule, unit, comment), encoding="utf-8")
