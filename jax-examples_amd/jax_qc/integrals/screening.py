"""Schwarz screening (replaces jax_qc/integrals/screening.py:32-65).

Q[i, j] = sqrt(max_ab |(ij|ij)|) over the Cartesian diagonal of the (ij|ij)
block, computed by the ERI kernels in their Schwarz mode; a quartet is
bounded by Q[i, j] * Q[k, l].
"""

from __future__ import annotations

import numpy as np

from jax_qc import _native as nat
from jax_qc.integrals import engine as eng


def compute_shell_pair_bounds(basis, pair_table=None) -> np.ndarray:
    pt = pair_table or eng.PairTable(basis, nat.device())
    q = eng.schwarz(pt)
    ns = pt.nshell
    Q = np.zeros((ns, ns))
    a, b = pt.sp_shells[:, 0], pt.sp_shells[:, 1]
    Q[a, b] = q
    Q[b, a] = q
    return Q


def estimate_eri_upper_bound(Q, i: int, j: int, k: int, l: int) -> float:
    return float(Q[i, j] * Q[k, l])
