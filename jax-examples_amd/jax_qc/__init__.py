"""jax_qc on MI355X: the McMurchie-Davidson integral and Fock-build path of
the jax_qc package re-implemented as HIP kernels for gfx950 (libqcmd.so),
behind the reference's jax_qc.integrals / jax_qc.scf / jax_qc.grad call
surface.  Matrices are FP64 torch tensors on the GPU."""

from jax_qc.basis import build_basis_set  # noqa: F401
from jax_qc.core.types import (  # noqa: F401
    BasisSet,
    CalcConfig,
    IntegralSet,
    Molecule,
    SCFResult,
    SCFState,
    Shell,
    make_molecule,
)
from jax_qc import autotune  # noqa: F401  (installs the measured per-class kernel choice)
from jax_qc.integrals import compute_integrals  # noqa: F401
from jax_qc.profiling import StageTimer  # noqa: F401
from jax_qc.scf import run_rhf, run_scf, run_uhf  # noqa: F401

__version__ = "0.1.0"
