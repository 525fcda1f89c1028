"""Geometry optimisation (reference surface: jax_qc/geomopt)."""

from jax_qc.geomopt.optimizer import OptResult, energy_and_gradient, optimize_bfgs  # noqa: F401


def optimize_geometry(mol, basis_name="sto-3g", method="rhf", max_steps=50, grad_tol=1e-4,
                      step_tol=1e-6, scf_conv=1e-10, max_step_size=0.3, verbose=1, **kw):
    """Dispatcher with the reference signature (geomopt/interface.py:20-60)."""
    return optimize_bfgs(mol, basis_name, method, max_steps, grad_tol, step_tol, scf_conv,
                         max_step_size, verbose, **kw)
