"""The primitive-pair records the device path builds with torch equal the
NumPy restatement of the Gaussian product theorem (p, P, P-A, P-B,
K = c_a c_b exp(-ab/p |AB|^2), alpha per primitive pair; the reference
recomputes them inside every quartet, obara_saika.py:56-65 /
gaussian_product.py:23-45), record for record, with and without the
primitive cut; runs on the CPU."""

import numpy as np
import torch

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from jax_qc.integrals import engine as eng


def _pairs(bs):
    sa = eng.shell_arrays(bs)
    L = sa["l"]
    ii, jj = np.tril_indices(len(bs.shells))
    swap = L[ii] < L[jj]
    return sa, np.where(swap, jj, ii).astype(np.int64), np.where(swap, ii, jj).astype(np.int64)


def test_device_records_equal_host_restatement():
    mol = M.water_cluster((2, 1, 1), seed=3)
    bs = build_basis_set(mol, "even-tempered")
    sa, pa, pb = _pairs(bs)
    for cut in (0.0, 1e-10):
        ref, cnt_ref = eng.all_pair_records(sa, pa, pb, cut)
        rec, cnt = eng.pair_records_device(sa, pa, pb, torch.device("cpu"), cut)
        np.testing.assert_array_equal(cnt, cnt_ref)
        np.testing.assert_allclose(rec.numpy(), ref, rtol=1e-14, atol=1e-300)
    # the host restatement itself against one shell pair computed directly
    k = 7
    one = eng.shell_pair_records(sa, int(pa[k]), int(pb[k]))
    ref, cnt_ref = eng.all_pair_records(sa, pa, pb)
    start = int(cnt_ref[:k].sum())
    np.testing.assert_allclose(ref[start:start + cnt_ref[k]], one, rtol=1e-15)
