"""Two-electron repulsion integrals (mu nu | lambda sigma) on the GPU.

``compute_eri_tensor`` keeps the reference's dense (n,n,n,n) result
(jax_qc/integrals/eri.py:69-105) for small bases and parity checks; large
bases use :class:`jax_qc.integrals.store.ERIStore` instead.
"""

from __future__ import annotations

import torch

from jax_qc import _native as nat
from jax_qc.integrals import engine as eng


def compute_eri_tensor(basis, tau: float = 0.0, tau32: float | None = None,
                       pair_table=None) -> torch.Tensor:
    """Dense ERI tensor, chemist's notation.  tau = 0 evaluates every
    canonical quartet, as the reference does."""
    dev = nat.device()
    nat.ensure_init(dev)
    n = int(basis.n_basis)
    if n ** 4 * 8 > 64 * 2 ** 30:
        raise ValueError(f"dense ERI tensor for n={n} needs {n**4*8/2**30:.0f} GiB; use ERIStore")
    pt = pair_table or eng.PairTable(basis, dev)
    plan = eng.plan_quartets(pt, tau, tau32)
    out = torch.zeros((n, n, n, n), dtype=torch.float64, device=dev)
    eng.run_eri(pt, plan, eng.MODE_DENSE, out)
    return out
