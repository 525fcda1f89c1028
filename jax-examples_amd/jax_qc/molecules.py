"""Seeded molecule generators for the benchmark workloads, plus the small
reference geometries the reference's tests and examples use.

All coordinates are returned in Bohr (ANGSTROM_TO_BOHR from
jax_qc/core/constants.py:14-15).  Every random draw comes from
``numpy.random.default_rng(seed)`` in a fixed order, so a (generator, seed)
pair names one geometry.
"""

from __future__ import annotations

import numpy as np

from jax_qc.core.constants import ANGSTROM_TO_BOHR, symbol_to_z
from jax_qc.core.types import Molecule, make_molecule

# Monomer geometry (Angstrom): O at the origin, H at (0, +-0.7572, 0.5865).
WATER_ANGSTROM = np.array([[0.0, 0.0, 0.0], [0.0, 0.7572, 0.5865], [0.0, -0.7572, 0.5865]])


def _random_rotation(rng) -> np.ndarray:
    """Uniform rotation from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


def water_cluster(grid=(2, 2, 2), seed: int = 0, spacing: float = 3.0, jitter: float = 0.2,
                  charge: int = 0, spin: int = 0) -> Molecule:
    """(H2O)_n on an nx x ny x nz grid (Angstrom spacing), each monomer
    randomly rotated about its O and jittered by U(-jitter, jitter) per atom
    coordinate.  Draw order per monomer (x slowest, z fastest): 4 normals for
    the rotation, then 9 uniforms for the jitter."""
    rng = np.random.default_rng(seed)
    elements, coords = [], []
    for ix in range(grid[0]):
        for iy in range(grid[1]):
            for iz in range(grid[2]):
                R = _random_rotation(rng)
                xyz = WATER_ANGSTROM @ R.T + spacing * np.array([ix, iy, iz], dtype=np.float64)
                xyz = xyz + rng.uniform(-jitter, jitter, size=(3, 3))
                elements += ["O", "H", "H"]
                coords.append(xyz)
    coords = np.concatenate(coords) * ANGSTROM_TO_BOHR
    Z = [symbol_to_z(e) for e in elements]
    return make_molecule(tuple(elements), coords, Z, charge=charge, spin=spin)


def water_monomer(angstrom: bool = True) -> Molecule:
    """The single water of the H2O STO-3G workload (Angstrom geometry)."""
    c = WATER_ANGSTROM * (ANGSTROM_TO_BOHR if angstrom else 1.0)
    return make_molecule(("O", "H", "H"), c, [8, 1, 1])


def random_molecule(rng, min_atoms: int = 2, max_atoms: int = 12) -> Molecule:
    """Random closed-shell {H,C,N,O} molecule.

    Heavy atoms (1..min(6, n-1)) grow as a random tree: each attaches to a
    random earlier heavy atom along a random direction at U(1.0, 1.5) A;
    hydrogens cap random heavy atoms at the same bond-length law.  Draws
    closer than 0.9 A to an existing atom are redrawn.  One hydrogen is
    added (or removed) to make the electron count even.
    """
    n_atoms = int(rng.integers(min_atoms, max_atoms + 1))
    n_heavy = int(rng.integers(1, min(6, n_atoms - 1) + 1)) if n_atoms > 1 else 1
    heavy_pool = ["C", "N", "O"]
    elements = [heavy_pool[int(rng.integers(0, 3))]]
    coords = [np.zeros(3)]

    def place(parent):
        for _ in range(200):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            pos = coords[parent] + d * rng.uniform(1.0, 1.5)
            if min(np.linalg.norm(pos - c) for c in coords) >= 0.9:
                return pos
        raise RuntimeError("could not place atom")

    for _ in range(1, n_heavy):
        parent = int(rng.integers(0, len(coords)))
        coords.append(place(parent))
        elements.append(heavy_pool[int(rng.integers(0, 3))])
    for _ in range(n_atoms - n_heavy):
        parent = int(rng.integers(0, n_heavy))
        coords.append(place(parent))
        elements.append("H")
    Z = [symbol_to_z(e) for e in elements]
    if sum(Z) % 2 == 1:
        if elements[-1] == "H" and len(elements) > 2:
            elements.pop()
            coords.pop()
        else:
            coords.append(place(int(rng.integers(0, n_heavy))))
            elements.append("H")
        Z = [symbol_to_z(e) for e in elements]
    c = np.array(coords) * ANGSTROM_TO_BOHR
    return make_molecule(tuple(elements), c, Z)


def random_molecules(n: int = 256, seed: int = 0, **kw):
    rng = np.random.default_rng(seed)
    return [random_molecule(rng, **kw) for _ in range(n)]


# --- small molecules from the reference's tests / examples (Bohr) ---------

def h2() -> Molecule:
    return make_molecule(("H", "H"), [[0, 0, 0], [0, 0, 1.4]], [1, 1])


def he() -> Molecule:
    return make_molecule(("He",), [[0, 0, 0]], [2])


def heh_cation() -> Molecule:
    return make_molecule(("He", "H"), [[0, 0, 0], [0, 0, 1.4632]], [2, 1], charge=1)


def h3_cation() -> Molecule:
    r = 1.65
    return make_molecule(("H", "H", "H"), [[0, 0, 0], [r, 0, 0], [r / 2, r * np.sqrt(3) / 2, 0]],
                         [1, 1, 1], charge=1)


def h2o_tier2() -> Molecule:
    """test_scf_tier2.py:26-37 geometry; the numbers are used as Bohr."""
    return make_molecule(("O", "H", "H"),
                         [[0, 0, 0.117], [0, 0.757, -0.469], [0, -0.757, -0.469]], [8, 1, 1])


def hf() -> Molecule:
    return make_molecule(("F", "H"), [[0, 0, 0], [0, 0, 1.7325]], [9, 1])


def ch4() -> Molecule:
    r = 2.052 / np.sqrt(3.0)
    return make_molecule(("C", "H", "H", "H", "H"),
                         [[0, 0, 0], [r, r, r], [r, -r, -r], [-r, r, -r], [-r, -r, r]],
                         [6, 1, 1, 1, 1])


def nh3() -> Molecule:
    """examples/08_rhf_tier2.py:83-97: N-H 1.913 Bohr, HNH 106.7 deg."""
    r = 1.913
    theta = np.deg2rad(106.7)
    s, c = np.sin(theta / 2.0), np.cos(theta / 2.0)
    return make_molecule(
        ("N", "H", "H", "H"),
        [[0, 0, 0], [r * s, 0, -r * c],
         [-r * s / 2, r * s * np.sqrt(3) / 2, -r * c],
         [-r * s / 2, -r * s * np.sqrt(3) / 2, -r * c]],
        [7, 1, 1, 1])


def oh_radical() -> Molecule:
    """Doublet OH at 1.8324 Bohr (the reference's UHF tests)."""
    return make_molecule(("O", "H"), [[0, 0, 0], [0, 0, 1.8324]], [8, 1], spin=1)


def o2_triplet() -> Molecule:
    return make_molecule(("O", "O"), [[0, 0, 0], [0, 0, 2.282]], [8, 8], spin=2)
