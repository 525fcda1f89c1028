"""Fock damping (replaces jax_qc/scf/damping.py:17-20)."""


def damp(F_new, F_old, alpha: float):
    return (1.0 - alpha) * F_new + alpha * F_old
