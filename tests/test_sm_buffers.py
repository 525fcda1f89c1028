"""The kept FP64 supermatrix buffers of jax_qc.integrals.store (host logic,
on CPU tensors): a buffer is handed out again only once nothing else refers
to it, zeroed when asked, at most SM_BUFFERS_MAX are kept, free buffers of
another size are dropped, and release_supermatrix_buffers() forgets them."""

import torch

from jax_qc.integrals import store as S


def test_buffers_are_reused_only_when_free():
    S._SM_BUFFERS.clear()
    dev = torch.device("cpu")
    a = S._sm_buffer(16, dev, zero=True)
    pa = a.data_ptr()
    a.fill_(3.0)
    b = S._sm_buffer(16, dev, zero=False)
    assert b.data_ptr() != pa  # a is still in use
    del a
    c = S._sm_buffer(16, dev, zero=True)
    assert c.data_ptr() == pa and torch.count_nonzero(c) == 0
    d = S._sm_buffer(16, dev, zero=False)  # both kept buffers in use: a fresh one, not kept
    assert d.data_ptr() not in (pa, b.data_ptr()) and len(S._SM_BUFFERS) == S.SM_BUFFERS_MAX
    del b, c, d
    e = S._sm_buffer(8, dev, zero=False)  # another size: the free ones of size 16 are dropped
    assert e.numel() == 8 and [t.numel() for t in S._SM_BUFFERS] == [8]
    S.release_supermatrix_buffers()
    assert S._SM_BUFFERS == []
