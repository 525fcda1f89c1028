"""SCF machinery on FP64 device tensors (reference surface: jax_qc/scf/__init__.py)."""

from jax_qc.scf.damping import damp  # noqa: F401
from jax_qc.scf.density import density_rhf, density_uhf  # noqa: F401
from jax_qc.scf.diis import DIISHistory, diis_extrapolate  # noqa: F401
from jax_qc.scf.energy import electronic_energy_rhf, electronic_energy_uhf  # noqa: F401
from jax_qc.scf.fock import build_fock_rhf, build_fock_uhf  # noqa: F401
from jax_qc.scf.guess import core_guess, core_guess_uhf  # noqa: F401
from jax_qc.scf.interface import run_scf  # noqa: F401
from jax_qc.scf.orthogonalize import canonical_orthogonalization, symmetric_orthogonalization  # noqa: F401
from jax_qc.scf.rhf import run_rhf, scf_step_rhf  # noqa: F401
from jax_qc.scf.uhf import run_uhf, scf_step_uhf  # noqa: F401
