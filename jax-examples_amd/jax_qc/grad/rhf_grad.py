"""rhf_gradient with the reference's signature (jax_qc/grad/rhf_grad.py:106).

The reference differentiates the integral matrices by central finite
differences (step_size) at fixed D and W; here the same Hellmann-Feynman +
Pulay expression is contracted analytically on the GPU, so ``step_size`` is
accepted for call compatibility and unused.
"""

from __future__ import annotations

from jax_qc.basis import build_basis_set
from jax_qc.grad.analytic import energy_weighted_density, scf_gradient  # noqa: F401


def rhf_gradient(mol, result, basis_name="sto-3g", step_size: float = 1e-4, basis=None,
                 tau: float = 1e-13):
    """(n_atoms, 3) gradient dE/dR in Eh/Bohr of a converged SCF result."""
    bs = basis if basis is not None else build_basis_set(mol, basis_name)
    return scf_gradient(mol, bs, result, tau)


uhf_gradient = rhf_gradient
