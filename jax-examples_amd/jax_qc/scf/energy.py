"""Electronic energies (replaces jax_qc/scf/energy.py:17-38)."""

import torch


def electronic_energy_rhf(D, H_core, F):
    return 0.5 * torch.sum(D * (H_core + F))


def electronic_energy_uhf(D_alpha, D_beta, H_core, F_alpha, F_beta):
    return 0.5 * torch.sum(D_alpha * (H_core + F_alpha)) + 0.5 * torch.sum(D_beta * (H_core + F_beta))
