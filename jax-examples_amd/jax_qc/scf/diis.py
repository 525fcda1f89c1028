"""Pulay DIIS (replaces jax_qc/scf/diis.py:26-199).

Error e = X^T (F D S - S D F) X; ring buffer of (F, e) evicting the oldest;
B-matrix solve on the host (size <= diis_space + 1) with a fallback to the
un-extrapolated Fock when the system is singular (diis.py:97-100, 126-128).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch


@dataclass
class DIISHistory:
    max_size: int
    focks: List[torch.Tensor] = field(default_factory=list)
    errors: List[torch.Tensor] = field(default_factory=list)
    # Gram matrix <e_i, e_j> of the stored errors, kept in step with the
    # ring: a push adds one row (a matrix-vector product) instead of a
    # (k x n^2)(n^2 x k) product, which BLAS runs as a single tile.
    gram: Optional[np.ndarray] = None
    # max |e| of the last push (both spins for UHF), read back with the
    # Gram row, so the drivers' convergence test needs no transfer of its own
    err_max: float = float("nan")

    @property
    def size(self) -> int:
        return len(self.focks)


def diis_history_init(n_basis: int, max_size: int) -> DIISHistory:
    return DIISHistory(max_size=max_size)


diis_history_init_uhf = diis_history_init


def diis_error(F, D, S, X):
    return X.T @ (F @ D @ S - S @ D @ F) @ X


def _push(h: DIISHistory, F, e) -> DIISHistory:
    """Append (F, e), evicting the oldest pair when full; the new Gram row
    and max |e| come back to the host in one transfer."""
    if h.size == h.max_size:
        h.focks.pop(0)
        h.errors.pop(0)
        if h.gram is not None:
            h.gram = h.gram[1:, 1:]
    h.focks.append(F)
    h.errors.append(e)
    v = e.reshape(-1)
    vals = torch.stack([torch.dot(x.reshape(-1), v) for x in h.errors] + [v.abs().max()])
    vals = vals.cpu().numpy()
    row, h.err_max = vals[:-1], float(vals[-1])
    m = len(h.errors)
    G = np.empty((m, m))
    if m > 1:
        G[:-1, :-1] = h.gram
    G[-1, :] = row
    G[:, -1] = row
    h.gram = G
    return h


def _solve_diis(errors, G=None) -> Optional[np.ndarray]:
    if G is None:
        E = torch.stack([e.reshape(-1) for e in errors])
        G = (E @ E.T).cpu().numpy()
    m = G.shape[0]
    B = np.empty((m + 1, m + 1))
    B[:m, :m] = G
    B[-1, :-1] = -1.0
    B[:-1, -1] = -1.0
    B[-1, -1] = 0.0
    rhs = np.zeros(m + 1)
    rhs[-1] = -1.0
    try:
        return np.linalg.solve(B, rhs)[:-1]
    except np.linalg.LinAlgError:
        return None


def _extrapolate(h: DIISHistory):
    c = _solve_diis(h.errors, h.gram)
    if c is None:
        return None
    ct = torch.as_tensor(c, dtype=h.focks[0].dtype, device=h.focks[0].device)
    return torch.einsum("i,imn->mn", ct, torch.stack(h.focks))


def diis_extrapolate(history: DIISHistory, F, D, S, X):
    e = diis_error(F, D, S, X)
    history = _push(history, F, e)
    if history.size < 2:
        return F, history, e
    Fx = _extrapolate(history)
    return (F if Fx is None else Fx), history, e


def diis_extrapolate_uhf(history: DIISHistory, F_a, F_b, D_a, D_b, S, X):
    n = F_a.shape[0]
    e_a = diis_error(F_a, D_a, S, X)
    e_b = diis_error(F_b, D_b, S, X)
    history = _push(history, torch.cat([F_a, F_b]), torch.cat([e_a, e_b]))
    if history.size < 2:
        return F_a, F_b, history, e_a, e_b
    Fx = _extrapolate(history)
    if Fx is None:
        return F_a, F_b, history, e_a, e_b
    return Fx[:n], Fx[n:], history, e_a, e_b
