"""DIIS bookkeeping on the host (CPU tensors): the Gram matrix kept in step
with the ring equals E E^T of the stored errors after every push, evictions
included (the reference recomputes B from the stored errors each time,
jax_qc/scf/diis.py:70-100)."""

import numpy as np
import torch

from jax_qc.scf.diis import _push, _solve_diis, diis_history_init


def test_incremental_gram_matches_full_product_through_evictions():
    g = torch.Generator().manual_seed(0)
    h = diis_history_init(6, 4)
    for it in range(9):
        e = torch.randn(6, 6, generator=g, dtype=torch.float64)
        F = torch.randn(6, 6, generator=g, dtype=torch.float64)
        h = _push(h, F, e)
        E = torch.stack([x.reshape(-1) for x in h.errors])
        np.testing.assert_allclose(h.gram, (E @ E.T).numpy(), rtol=1e-13, atol=1e-12)
        assert h.size == min(it + 1, 4)
        if h.size >= 2:
            np.testing.assert_allclose(_solve_diis(h.errors, h.gram), _solve_diis(h.errors),
                                       rtol=1e-10, atol=1e-12)
