"""Orthogonalisers (replaces jax_qc/scf/orthogonalize.py:22-41)."""

import torch


def symmetric_orthogonalization(S):
    w, U = torch.linalg.eigh(S)
    return (U * (1.0 / torch.sqrt(w))[None, :]) @ U.T


def canonical_orthogonalization(S, eps: float = 1e-7):
    w, U = torch.linalg.eigh(S)
    keep = w > eps
    return U[:, keep] / torch.sqrt(w[keep])[None, :]
