"""Boys function F_n(t) on the GPU (replaces jax_qc/integrals/boys.py:58-109).

Unlike the reference's upward recursion (accurate only for small n away
from t ~ 0.01), the device kernel uses pretabulated Taylor expansion plus
downward recursion and is accurate to ~1e-15 relative for every n <= 12.
"""

from __future__ import annotations

import numpy as np
import torch

from jax_qc import _native as nat


def boys_table(nmax: int, t, dtype=torch.float64) -> torch.Tensor:
    """F_0..F_nmax at every t: shape t.shape + (nmax+1,)."""
    if nmax < 0:
        raise ValueError("boys_fn requires n >= 0")
    dev = nat.device()
    nat.ensure_init(dev)
    tt = torch.as_tensor(np.asarray(t, dtype=np.float64) if not isinstance(t, torch.Tensor) else t)
    shape = tuple(tt.shape)
    tt = tt.reshape(-1).to(device=dev, dtype=dtype).contiguous()
    F = torch.empty(tt.numel() * (nmax + 1), dtype=dtype, device=dev)
    fn = "qc_boys" if dtype == torch.float64 else "qc_boys_f32"
    nat.call(fn, nmax, nat.ptr(tt), nat.ptr(F), tt.numel(), nat.stream_ptr())
    return F.reshape(*shape, nmax + 1)


def boys_fn(n: int, t) -> torch.Tensor:
    if n < 0:
        raise ValueError("boys_fn requires n >= 0")
    return boys_table(n, t)[..., n]


def boys_f0(t) -> torch.Tensor:
    return boys_table(0, t)[..., 0]
