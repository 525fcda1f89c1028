"""Analytic and numerical nuclear gradients (reference surface: jax_qc/grad)."""

from jax_qc.grad.analytic import energy_weighted_density, scf_gradient  # noqa: F401
from jax_qc.grad.autograd import SCFEnergy, scf_energy  # noqa: F401
from jax_qc.grad.interface import compute_gradient  # noqa: F401
from jax_qc.grad.numerical_grad import numerical_gradient  # noqa: F401
from jax_qc.grad.rhf_grad import rhf_gradient, uhf_gradient  # noqa: F401
