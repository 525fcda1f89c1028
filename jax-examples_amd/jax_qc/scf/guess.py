"""Initial guesses (replaces jax_qc/scf/guess.py:25-115)."""

import torch

from jax_qc.scf import purify
from jax_qc.scf.density import density_rhf
from jax_qc.scf.fock import _coulomb, _exchange


def _diag(F, X):
    e, Cp = torch.linalg.eigh(X.T @ F @ X)
    return e, X @ Cp


def _occupied_density(F, X, n_occ: int):
    """C_occ C_occ^T of the n_occ lowest orbitals of F, by purification
    (jax_qc.scf.purify: the same projector as the eigenvectors give, with
    matrix products only); None when purification is off or does not
    converge."""
    if not purify.ENABLED:
        return None
    Xt = X.T.contiguous()
    P = purify.occupied_projectors((Xt @ F @ X)[None], [n_occ])
    return None if P is None else X @ P[0] @ Xt


def core_guess(H_core, X, n_occ: int):
    D = _occupied_density(H_core, X, n_occ)
    if D is not None:
        return 2.0 * D
    return density_rhf(_diag(H_core, X)[1], n_occ)


def core_guess_uhf(H_core, X, n_alpha: int, n_beta: int, ERI, orbitals: bool = True):
    """Alpha from the core Hamiltonian; for open shells beta from
    H + J(Da) - K(Da) plus the deterministic sin/cos perturbation scaled by
    1% of the eigenvalue spread + 0.1 (guess.py:87-106).  Returns (Da, Db,
    alpha orbitals); orbitals=False takes both densities from purification
    where it converges (the SCF drivers need only the densities) and returns
    None for the orbitals then."""
    Ca = None
    Da = None if orbitals else _occupied_density(H_core, X, n_alpha)
    if Da is None:
        _, Ca = _diag(H_core, X)
        Da = Ca[:, :n_alpha] @ Ca[:, :n_alpha].T
    if n_alpha == n_beta:
        return Da, Da.clone(), Ca
    Fb = H_core + _coulomb(Da, ERI) - _exchange(Da, ERI)
    ev = torch.linalg.eigvalsh(X.T @ Fb @ X)
    scale = 0.01 * float(ev.max() - ev.min()) + 0.1
    idx = torch.arange(H_core.shape[0], dtype=torch.float64, device=H_core.device)
    noise = scale * torch.outer(torch.sin(idx * 1.23 + 0.45), torch.cos(idx * 0.67 + 0.89))
    Fb = Fb + (noise + noise.T) / 2.0
    Db = None if orbitals else _occupied_density(Fb, X, n_beta)
    if Db is None:
        _, Cb = _diag(Fb, X)
        Db = Cb[:, :n_beta] @ Cb[:, :n_beta].T
    return Da, Db, Ca
