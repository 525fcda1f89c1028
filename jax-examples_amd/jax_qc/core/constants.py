"""Unit conversions and element data (atomic units internally).

Mirrors jax_qc/core/constants.py:12-70 of the reference: CODATA 2018 Bohr
radius, element symbol table and symbol<->Z helpers.
"""

from __future__ import annotations

BOHR_TO_ANGSTROM: float = 0.529177210903
ANGSTROM_TO_BOHR: float = 1.0 / BOHR_TO_ANGSTROM

HARTREE_TO_EV: float = 27.211386245988
HARTREE_TO_KCAL_PER_MOL: float = 627.5094740631

ELEMENT_SYMBOLS = (
    "X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
    "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar",
)
SYMBOL_TO_Z = {s: z for z, s in enumerate(ELEMENT_SYMBOLS)}


def symbol_to_z(symbol: str) -> int:
    key = symbol.strip().title()
    if key not in SYMBOL_TO_Z:
        raise ValueError(f"Unknown element symbol: {symbol!r}")
    return SYMBOL_TO_Z[key]


def z_to_symbol(z: int) -> str:
    if z < 0 or z >= len(ELEMENT_SYMBOLS):
        raise ValueError(f"Atomic number out of range: {z}")
    return ELEMENT_SYMBOLS[z]
