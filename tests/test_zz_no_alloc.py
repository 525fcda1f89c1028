"""qc_symv_noalloc (the Fock-build SYMV with the library's static device
scratch instead of a caller workspace) against qc_symv on the same
supermatrix and vectors, on two calls in a row (the second on a smaller
input, so stale scratch contents from the first would show), and its
size limit reported as an error before any launch."""

import ctypes

import numpy as np
import pytest
import torch

from jax_qc import _native as nat
from jax_qc import molecules as M
from jax_qc.basis import build_basis_set

pytestmark = pytest.mark.gpu


def _symv_both(mol, nrhs, seed):
    from jax_qc.integrals.store import ERIStore

    bs = build_basis_set(mol, "even-tempered")
    st = ERIStore(bs, tau=1e-12)
    P = st.ensure_P()
    n = bs.n_basis
    g = torch.Generator(device="cpu").manual_seed(seed)
    A = torch.randn(nrhs, n, n, generator=g, dtype=torch.float64)
    D = (A + A.transpose(1, 2)).cuda().contiguous()
    work, x, _ = st._buffers(nrhs)
    s = nat.stream_ptr()
    nat.call("qc_pack_density", nat.ptr(D), n, nrhs, nat.ptr(x), s)
    y_ref = torch.empty_like(x)
    y_na = torch.full_like(x, np.nan)
    nat.call("qc_symv", nat.ptr(P), n, nrhs, nat.ptr(x), nat.ptr(y_ref), nat.ptr(work), s)
    nat.call("qc_symv_noalloc", nat.ptr(P), n, nrhs, nat.ptr(x), nat.ptr(y_na), s)
    torch.cuda.synchronize()
    return y_ref, y_na


def test_noalloc_symv_equals_workspace_symv_twice_second_smaller():
    print("[case] qc_symv_noalloc: (H2O)_2 then H2O", flush=True)
    nat.ensure_init()
    for mol, nrhs, seed in ((M.water_cluster((2, 1, 1), seed=4), 2, 0),
                            (M.water_cluster((1, 1, 1), seed=4), 1, 1)):
        y_ref, y_na = _symv_both(mol, nrhs, seed)
        assert torch.isfinite(y_na).all()
        assert torch.equal(y_ref, y_na)


def test_noalloc_symv_rejects_sizes_beyond_its_scratch():
    nat.ensure_init()
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    rc = nat.lib().qc_symv_noalloc(nat.ptr(dummy), 4000, 1, nat.ptr(dummy), nat.ptr(dummy),
                                   nat.stream_ptr())
    assert rc != 0
    msg = nat.lib().qc_last_error().decode()
    assert "scratch" in msg and "4000" in msg
