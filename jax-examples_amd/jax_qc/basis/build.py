"""Basis construction: BSE-layout dict -> ordered, normalised shells.

Follows the reference's conventions exactly, because they fix the AO order
and the contraction coefficients of every output:

* one Shell per (angular momentum, coefficient row); combined sp shells are
  split (bse_fetch.py:35-40 ``uncontract_spdf=True``; parse.py:26-89);
* primitives with an exactly-zero coefficient are dropped (parse.py:70-78);
* per atom, shells are sorted by (l ascending, max exponent descending)
  (parse.py:120-125);
* spherical (2l+1) layout by default (build.py:28).

``build_basis_set`` accepts a library name ("sto-3g", "6-31g", "6-31g**"),
"even-tempered" (the synthetic benchmark basis, see ``even_tempered_dict``)
or a BSE JSON dict.
"""

from __future__ import annotations

import json
from typing import Dict, List, Union

import numpy as np

from jax_qc.basis.library import get_basis_dict
from jax_qc.basis.normalize import normalize_shell
from jax_qc.core.types import BasisSet, Molecule, Shell

# Synthetic even-tempered basis: alpha_k = ALPHA0 * RATIO**k, each shell a
# 3-primitive contraction over consecutive k.  Exponent-index triplets per
# shell, per element class.
ALPHA0 = 0.1
RATIO = 2.5
_ET_LAYOUT = {
    "H": {0: [(0, 1, 2), (2, 3, 4), (4, 5, 6)], 1: [(1, 2, 3)]},
    "heavy": {
        0: [(0, 1, 2), (2, 3, 4), (4, 5, 6), (6, 7, 8)],
        1: [(0, 1, 2), (2, 3, 4), (4, 5, 6)],
        2: [(1, 2, 3)],
    },
}
ET_BASIS_SEED = 0


def even_tempered_dict(seed: int = ET_BASIS_SEED, elements=(1, 6, 7, 8)) -> Dict:
    """BSE-layout dict of the synthetic even-tempered basis.

    H: 3s1p; C, N, O: 4s3p1d.  Raw contraction coefficients are drawn from
    ``numpy.random.default_rng(seed)`` as U(0.25, 1.0), element by element in
    ascending Z, shell by shell in (l, shell) order, then normalised exactly
    as any other basis.
    """
    rng = np.random.default_rng(seed)
    out = {"name": "even-tempered", "elements": {}}
    for z in sorted(elements):
        layout = _ET_LAYOUT["H" if z == 1 else "heavy"]
        shells = []
        for l in sorted(layout):
            for ks in layout[l]:
                exps = [ALPHA0 * RATIO**k for k in ks]
                coefs = rng.uniform(0.25, 1.0, size=len(ks))
                shells.append(
                    {
                        "angular_momentum": [l],
                        "exponents": [repr(float(e)) for e in exps],
                        "coefficients": [[repr(float(c)) for c in coefs]],
                    }
                )
        out["elements"][str(z)] = {"electron_shells": shells}
    return out


def parse_electron_shell(shell_dict: Dict, center: np.ndarray, atom_index: int) -> List[Shell]:
    """One BSE electron_shell entry -> Shells (parse.py:26-89)."""
    ams = list(shell_dict["angular_momentum"])
    exps = np.array([float(s) for s in shell_dict["exponents"]], dtype=np.float64)
    rows = shell_dict["coefficients"]
    if len(ams) == len(rows):
        pairs = list(zip(ams, rows))
    elif len(ams) == 1:
        pairs = [(ams[0], r) for r in rows]
    else:
        raise ValueError(f"Unexpected BSE shell layout: {len(ams)} angular momenta, {len(rows)} rows.")
    out = []
    for l, row in pairs:
        c = np.array([float(s) for s in row], dtype=np.float64)
        mask = c != 0.0
        a, c = exps[mask], c[mask]
        out.append(
            Shell(
                angular_momentum=int(l),
                exponents=a,
                coefficients=normalize_shell(a, c, int(l)),
                center=np.asarray(center, dtype=np.float64).copy(),
                atom_index=int(atom_index),
            )
        )
    return out


def bse_dict_to_shells(bse: Dict, mol: Molecule) -> List[Shell]:
    """Per atom in order; within an atom sort by (l, -max exponent)
    (parse.py:92-127)."""
    elements = bse.get("elements", {})
    shells: List[Shell] = []
    for idx, z in enumerate(np.asarray(mol.atomic_numbers).tolist()):
        key = str(int(z))
        if key not in elements:
            raise KeyError(f"Basis set does not define element Z={z} (atom index {idx}).")
        atom = []
        for sd in elements[key].get("electron_shells", []):
            atom.extend(parse_electron_shell(sd, mol.coords[idx], idx))
        atom.sort(key=lambda sh: (sh.angular_momentum, -float(np.max(sh.exponents))))
        shells.extend(atom)
    return shells


def build_shell_indices(shells, spherical: bool = True):
    """(shell_to_basis, basis_to_atom, n_basis) (parse.py:130-152)."""
    s2b, b2a, off = [], [], 0
    for sh in shells:
        l = sh.angular_momentum
        n = (2 * l + 1) if spherical else (l + 1) * (l + 2) // 2
        s2b.append(tuple(range(off, off + n)))
        b2a.extend([sh.atom_index] * n)
        off += n
    return tuple(s2b), np.asarray(b2a, dtype=np.int32), off


def build_basis_set(
    mol: Molecule, basis_name: Union[str, Dict] = "sto-3g", spherical: bool = True
) -> BasisSet:
    """Build the normalised, ordered BasisSet for ``mol`` (build.py:24-62)."""
    if isinstance(basis_name, dict):
        bse, name = basis_name, basis_name.get("name", "custom")
    elif basis_name.strip().lower() == "even-tempered":
        bse, name = even_tempered_dict(), "even-tempered"
    elif basis_name.strip().endswith(".json"):
        with open(basis_name) as fh:
            bse = json.load(fh)
        name = bse.get("name", basis_name)
    else:
        bse, name = get_basis_dict(basis_name), basis_name.strip().lower()
    shells = bse_dict_to_shells(bse, mol)
    s2b, b2a, n = build_shell_indices(shells, spherical)
    return BasisSet(
        shells=tuple(shells),
        n_basis=n,
        shell_to_basis=s2b,
        basis_to_atom=b2a,
        name=name,
        spherical=spherical,
    )


def shell_arrays(basis: BasisSet):
    """Flatten a BasisSet into the plain arrays the C-ABI takes.

    Returns dict with l (int32), nprim (int32), prim_off (int32), exps,
    coefs (float64), centers (nshell, 3), atom (int32), ao_off (int32,
    nshell+1, in the basis' own layout).
    """
    shells = basis.shells
    l = np.array([s.angular_momentum for s in shells], dtype=np.int32)
    nprim = np.array([len(s.exponents) for s in shells], dtype=np.int32)
    prim_off = np.zeros(len(shells), dtype=np.int32)
    if len(shells) > 1:
        prim_off[1:] = np.cumsum(nprim)[:-1]
    exps = np.concatenate([s.exponents for s in shells]).astype(np.float64)
    coefs = np.concatenate([s.coefficients for s in shells]).astype(np.float64)
    centers = np.array([s.center for s in shells], dtype=np.float64).reshape(-1, 3)
    atom = np.array([s.atom_index for s in shells], dtype=np.int32)
    nf = (2 * l + 1) if basis.spherical else (l + 1) * (l + 2) // 2
    ao_off = np.zeros(len(shells) + 1, dtype=np.int32)
    ao_off[1:] = np.cumsum(nf)
    return dict(l=l, nprim=nprim, prim_off=prim_off, exps=exps, coefs=coefs,
                centers=centers, atom=atom, ao_off=ao_off)
