"""Gradient dispatcher (replaces jax_qc/grad/interface.py:20-55)."""

from jax_qc.grad.numerical_grad import numerical_gradient
from jax_qc.grad.rhf_grad import rhf_gradient


def compute_gradient(mol, result, basis_name="sto-3g", method="analytic", step_size=1e-4):
    if method == "analytic":
        return rhf_gradient(mol, result, basis_name=basis_name, step_size=step_size)
    if method == "numerical":
        return numerical_gradient(mol, basis_name=basis_name,
                                  method="rhf" if mol.spin == 0 else "uhf", step_size=step_size)
    raise ValueError(f"Unknown gradient method '{method}'; expected 'analytic' or 'numerical'.")
