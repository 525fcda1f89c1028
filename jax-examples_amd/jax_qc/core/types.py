"""Immutable carriers for the integral / Fock-build path.

Field names and meanings follow jax_qc/core/types.py:40-201 of the reference
(Shell, BasisSet, Molecule, IntegralSet, SCFState, SCFResult, CalcConfig,
make_molecule).  Geometry and basis data are host numpy arrays (float64,
Bohr); matrices produced by the device path are torch float64 tensors on the
GPU.  ``IntegralSet.ERI`` and ``SCFResult.ERI`` may hold either a dense
(n, n, n, n) tensor or an ERI store handle
(:class:`jax_qc.integrals.store.ERIStore`), the integral-direct hook the
reference leaves as ``Optional`` (types.py:150).
"""

from __future__ import annotations

import dataclasses
from typing import Any, Optional, Tuple

import numpy as np


@dataclasses.dataclass(frozen=True)
class Shell:
    """A contracted shell; ``coefficients`` already include primitive and
    contraction normalisation (types.py:40-58)."""

    angular_momentum: int
    exponents: np.ndarray
    coefficients: np.ndarray
    center: np.ndarray
    atom_index: int


@dataclasses.dataclass(frozen=True)
class BasisSet:
    """Shells ordered by atom, then (l ascending, max exponent descending)
    within an atom (types.py:61-81)."""

    shells: tuple
    n_basis: int
    shell_to_basis: tuple
    basis_to_atom: np.ndarray
    name: str = ""
    spherical: bool = True


@dataclasses.dataclass(frozen=True)
class Molecule:
    """Coordinates in Bohr (types.py:84-105)."""

    coords: np.ndarray
    atomic_numbers: np.ndarray
    elements: Tuple[str, ...]
    charge: int = 0
    spin: int = 0
    n_electrons: int = 0
    n_alpha: int = 0
    n_beta: int = 0


@dataclasses.dataclass(frozen=True)
class IntegralSet:
    S: Any
    T: Any
    V: Any
    H_core: Any
    ERI: Any
    E_nuc: float


@dataclasses.dataclass(frozen=True)
class SCFState:
    density: Any
    fock: Any
    coefficients: Any
    orbital_energies: Any
    energy: float
    iteration: int


@dataclasses.dataclass(frozen=True)
class SCFResult:
    converged: bool
    state: SCFState
    energy: float
    E_elec: float
    E_nuc: float
    n_iterations: int
    S: Any
    H_core: Any
    ERI: Optional[Any] = None
    # Present for UHF results (absent in the reference, which cannot form a
    # correct UHF gradient without them).
    uhf: Optional[Any] = None


@dataclasses.dataclass(frozen=True)
class CalcConfig:
    method: str = "rhf"
    basis: str = "sto-3g"
    task: str = "energy"
    max_scf_iter: int = 128
    scf_conv: float = 1e-10
    diis_space: int = 8
    damping: float = 0.0
    guess: str = "core"
    verbose: int = 1
    profile: bool = False


def make_molecule(elements, coords, atomic_numbers, charge: int = 0, spin: int = 0) -> Molecule:
    """Build a Molecule and derive electron counts (types.py:169-201)."""
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    Z = np.asarray(atomic_numbers, dtype=np.int64).reshape(-1)
    n_electrons = int(Z.sum()) - int(charge)
    if n_electrons < 0:
        raise ValueError(f"Negative electron count: charge {charge} exceeds nuclear charge.")
    if (n_electrons - int(spin)) % 2 != 0:
        raise ValueError(
            f"Inconsistent spin: n_electrons={n_electrons}, 2S={spin} have different parity."
        )
    n_beta = (n_electrons - int(spin)) // 2
    return Molecule(
        coords=coords,
        atomic_numbers=Z,
        elements=tuple(elements),
        charge=int(charge),
        spin=int(spin),
        n_electrons=n_electrons,
        n_alpha=n_electrons - n_beta,
        n_beta=n_beta,
    )
