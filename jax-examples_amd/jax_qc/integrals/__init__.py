"""Molecular integrals on the GPU (reference surface: jax_qc/integrals/__init__.py)."""

from jax_qc.integrals.boys import boys_f0, boys_fn  # noqa: F401
from jax_qc.integrals.eri import compute_eri_tensor  # noqa: F401
from jax_qc.integrals.interface import compute_integrals  # noqa: F401
from jax_qc.integrals.one_electron import (  # noqa: F401
    compute_kinetic_matrix,
    compute_nuclear_matrix,
    compute_overlap_matrix,
    nuclear_repulsion_energy,
)
from jax_qc.integrals.screening import compute_shell_pair_bounds, estimate_eri_upper_bound  # noqa: F401
from jax_qc.integrals.store import ERIStore  # noqa: F401
