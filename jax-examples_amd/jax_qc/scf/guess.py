"""Initial guesses (replaces jax_qc/scf/guess.py:25-115)."""

import torch

from jax_qc.scf.density import density_rhf, density_uhf
from jax_qc.scf.fock import _coulomb, _exchange


def _diag(F, X):
    e, Cp = torch.linalg.eigh(X.T @ F @ X)
    return e, X @ Cp


def core_guess(H_core, X, n_occ: int):
    return density_rhf(_diag(H_core, X)[1], n_occ)


def core_guess_uhf(H_core, X, n_alpha: int, n_beta: int, ERI):
    """Alpha from the core Hamiltonian; for open shells beta from
    H + J(Da) - K(Da) plus the deterministic sin/cos perturbation scaled by
    1% of the eigenvalue spread + 0.1 (guess.py:87-106)."""
    _, Ca = _diag(H_core, X)
    if n_alpha != n_beta:
        Da0 = Ca[:, :n_alpha] @ Ca[:, :n_alpha].T
        Fb = H_core + _coulomb(Da0, ERI) - _exchange(Da0, ERI)
        ev = torch.linalg.eigvalsh(X.T @ Fb @ X)
        scale = 0.01 * float(ev.max() - ev.min()) + 0.1
        idx = torch.arange(H_core.shape[0], dtype=torch.float64, device=H_core.device)
        noise = scale * torch.outer(torch.sin(idx * 1.23 + 0.45), torch.cos(idx * 0.67 + 0.89))
        Fb = Fb + (noise + noise.T) / 2.0
        _, Cb = _diag(Fb, X)
    else:
        Cb = Ca
    Da, Db = density_uhf(Ca, Cb, n_alpha, n_beta)
    return Da, Db, Ca
