"""Fock assembly (replaces jax_qc/scf/fock.py:31-79).

``ERI`` is either a dense (n,n,n,n) device tensor, contracted by the
qc_dense_jk kernel with the reference's einsum semantics
(J_mn = sum_ls D_ls (mn|ls), K_mn = sum_ls D_ls (ms|ln)), or an
:class:`ERIStore`, contracted by one supermatrix-vector product per build.
"""

from __future__ import annotations

import torch

from jax_qc import _native as nat
from jax_qc.integrals.store import ERIStore


def _dense_jk(D, ERI):
    n = D.shape[0]
    J = torch.empty_like(D)
    K = torch.empty_like(D)
    nat.call("qc_dense_jk", nat.ptr(ERI.contiguous()), nat.ptr(D.contiguous()), n, nat.ptr(J),
             nat.ptr(K), nat.stream_ptr())
    return J, K


def _coulomb(D, ERI):
    if isinstance(ERI, ERIStore):
        return ERI.coulomb(D)
    return _dense_jk(D, ERI)[0]


def _exchange(D, ERI):
    if isinstance(ERI, ERIStore):
        return ERI.exchange(D)
    return _dense_jk(D, ERI)[1]


def build_fock_rhf(H_core, D, ERI):
    """F = H + J(D) - 1/2 K(D), D the total density."""
    if isinstance(ERI, ERIStore):
        return ERI.fock_rhf(H_core, D)
    J, K = _dense_jk(D, ERI)
    return H_core + J - 0.5 * K


def build_fock_uhf(H_core, D_alpha, D_beta, ERI):
    """F_s = H + J(D_a + D_b) - K(D_s)."""
    if isinstance(ERI, ERIStore):
        return ERI.fock_uhf(H_core, D_alpha, D_beta)
    J, _ = _dense_jk(D_alpha + D_beta, ERI)
    _, Ka = _dense_jk(D_alpha, ERI)
    _, Kb = _dense_jk(D_beta, ERI)
    return H_core + J - Ka, H_core + J - Kb


def last_build_was_full(ERI) -> bool:
    """True unless ERI is an ERIStore in incremental mode whose last Fock
    build was incremental; in that case the next build is made in full (the
    SCF drivers call this before accepting convergence, so a converged
    result always rests on a full build)."""
    if isinstance(ERI, ERIStore) and ERI.last_build_incremental():
        ERI.request_full_build()
        return False
    return True
