"""Write SCF / gradient fixtures computed by the CPU oracle (oracle/, the
restatement of the reference's integrals, SCF and gradient), for the GPU
parity tests of configurations too large for the oracle to run inside a
test:

* ``water4_rhf_grad.json``: (H2O)_4, grid 2x2x1, seed 1, even-tempered
  basis (120 functions), RHF at conv 1e-12 and the reference's gradient
  algorithm -- central finite differences (h = 1e-4 Bohr) of the integrals
  at fixed D and W (jax_qc/grad/rhf_grad.py:106-211) on all 36 coordinates.
  The displaced ERI rebuilds compute the quartets that touch the displaced
  atom only (oracle.scf.rhf_gradient(touching_only=True)); every other
  element has a zero difference.
* ``water2_uhf_triplet.json``: (H2O)_2, grid 2x1x1, seed 1, spin 2,
  even-tempered basis (60 functions), UHF at conv 1e-12
  (jax_qc/scf/uhf.py:187-287).
* ``batch6_rhf.json``: the six molecules molecules.random_molecules(6,
  seed=0, max_atoms=6) with the even-tempered basis, RHF at conv 1e-10
  (jax_qc/scf/rhf.py:52-169).

Each file also holds the input coordinates, so a test can check that the
generators still produce the same molecules.

Run: python tests/golden/make_scf_golden.py [water4|water2|batch6 ...] [--cache=DIR]
(the water4 gradient takes about an hour on 8 cores; --cache keeps per-atom
rows so an interrupted run resumes).
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for _p in (REPO, os.path.join(REPO, "jax-examples_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

from jax_qc import molecules as M  # noqa: E402
from jax_qc.basis import build_basis_set  # noqa: E402
from jax_qc.core.types import make_molecule  # noqa: E402
from oracle import mmd, scf  # noqa: E402


def _displace(mol, c):
    return make_molecule(mol.elements, c, mol.atomic_numbers, mol.charge, mol.spin)


def _bf(m):
    return build_basis_set(m, "even-tempered")


def _mol_record(mol):
    return {"elements": list(mol.elements), "coords_bohr": np.asarray(mol.coords).tolist(),
            "charge": int(mol.charge), "spin": int(mol.spin)}


def _write(name, obj):
    path = os.path.join(HERE, name)
    with open(path, "w") as fh:
        json.dump(obj, fh, indent=1)
    print("wrote", path, flush=True)


def water4(cache=None):
    """cache: a directory for per-atom rows, so an interrupted run resumes."""
    mol = M.water_cluster((2, 2, 1), seed=1)
    bs = _bf(mol)
    t = time.time()
    res = scf.run_rhf(mol, mmd.compute_integrals(mol, bs), conv=1e-12)
    assert res["converged"]
    g = np.zeros((len(mol.atomic_numbers), 3))
    for a in range(len(mol.atomic_numbers)):
        path = os.path.join(cache, f"water4_atom{a}.json") if cache else None
        if path and os.path.exists(path):
            with open(path) as fh:
                g[a] = json.load(fh)
            continue
        g[a] = scf.rhf_gradient(mol, res, _bf, _displace, step=1e-4, touching_only=True,
                                atoms=[a])[a]
        print(f"atom {a}: {g[a].tolist()}  ({time.time() - t:.0f} s)", flush=True)
        if path:
            os.makedirs(cache, exist_ok=True)
            with open(path, "w") as fh:
                json.dump(g[a].tolist(), fh)
    _write("water4_rhf_grad.json", {
        "molecule": _mol_record(mol), "basis": "even-tempered", "n_basis": bs.n_basis,
        "scf_conv": 1e-12, "energy": res["energy"], "n_iter": res["n_iter"], "fd_step": 1e-4,
        "gradient": g.tolist(), "seconds": time.time() - t})


def water2():
    mol = M.water_cluster((2, 1, 1), seed=1, spin=2)
    bs = _bf(mol)
    res = scf.run_uhf(mol, mmd.compute_integrals(mol, bs), conv=1e-12)
    assert res["converged"]
    _write("water2_uhf_triplet.json", {
        "molecule": _mol_record(mol), "basis": "even-tempered", "n_basis": bs.n_basis,
        "scf_conv": 1e-12, "energy": res["energy"], "n_iter": res["n_iter"]})


def batch6():
    mols = M.random_molecules(6, seed=0, max_atoms=6)
    out = []
    for m in mols:
        bs = _bf(m)
        res = scf.run_rhf(m, mmd.compute_integrals(m, bs), conv=1e-10)
        out.append({"molecule": _mol_record(m), "n_basis": bs.n_basis, "energy": res["energy"],
                    "converged": bool(res["converged"]), "n_iter": res["n_iter"]})
    _write("batch6_rhf.json", {"generator": "random_molecules(6, seed=0, max_atoms=6)",
                               "basis": "even-tempered", "scf_conv": 1e-10, "molecules": out})


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--cache=")]
    cache = next((a.split("=", 1)[1] for a in sys.argv[1:] if a.startswith("--cache=")), None)
    which = args or ["water2", "batch6", "water4"]
    for w in which:
        if w == "water4":
            water4(cache)
        else:
            {"water2": water2, "batch6": batch6}[w]()
