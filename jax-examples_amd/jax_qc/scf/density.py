"""Densities (replaces jax_qc/scf/density.py:23-58)."""


def density_rhf(C, n_occ: int):
    Co = C[:, :n_occ]
    return 2.0 * Co @ Co.T


def density_uhf(C_alpha, C_beta, n_alpha: int, n_beta: int):
    Ca, Cb = C_alpha[:, :n_alpha], C_beta[:, :n_beta]
    return Ca @ Ca.T, Cb @ Cb.T
