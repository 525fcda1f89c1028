"""SCF dispatcher (replaces jax_qc/scf/interface.py:17-31)."""

from jax_qc.scf.rhf import run_rhf
from jax_qc.scf.uhf import run_uhf


def run_scf(mol, ints, config, timer=None):
    method = config.method.lower()
    if method == "rhf":
        return run_rhf(mol, ints, config, timer=timer)
    if method == "uhf":
        return run_uhf(mol, ints, config, timer=timer)
    raise NotImplementedError(f"SCF method '{config.method}' is not implemented; available: rhf, uhf.")
