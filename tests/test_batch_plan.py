"""Batched quartet planning (CPU only): the plan for many molecules must hold
exactly the union of the single-molecule plans (the reference's canonical
quartets i>=j, k>=l, ij>=kl per molecule, eri.py:89-104, after the same
Schwarz screening), and no quartet may pair shell pairs of two molecules."""

import numpy as np
import pytest
import torch

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from jax_qc.batch import BatchPairTable, plan_quartets_batched
from jax_qc.integrals import engine as eng


def _quartets(plan):
    out = set()
    for L in plan.launches:
        bra, ket = L.bra.numpy(), L.ket.numpy()
        offs, ks = L.offsets.numpy(), L.kstart.numpy()
        for b in range(L.nbra):
            for j in range(offs[b + 1] - offs[b]):
                out.add((int(bra[b]), int(ket[ks[b] + j]), L.fp32))
    return out


@pytest.mark.parametrize("tau,tau32", [(0.0, None), (1e-6, None), (1e-8, 1e-4)])
def test_batched_plan_is_union_of_single_plans(tau, tau32):
    cpu = torch.device("cpu")
    rng = np.random.default_rng(3)
    mols = M.random_molecules(4, seed=7, max_atoms=4)
    tables = [eng.PairTable(build_basis_set(m, "sto-3g"), cpu) for m in mols]
    for t in tables:
        t.Q = rng.uniform(1e-6, 1.0, t.npairs) ** 3  # synthetic bounds, spread over decades
    bpt = BatchPairTable(tables, cpu)
    bpt.Q = np.concatenate([t.Q for t in tables])
    got = _quartets(plan_quartets_batched(bpt, tau, tau32))
    want = set()
    p0 = 0
    for t in tables:
        for b, k, f in _quartets(eng.plan_quartets(t, tau, tau32)):
            want.add((b + p0, k + p0, f))
        p0 += t.npairs
    assert got == want
    mol = bpt.sp_mol
    assert all(mol[b] == mol[k] for b, k, _ in got)
    if tau == 0.0:  # unscreened: every canonical quartet of every molecule, once
        n = sum(len(t.sp_class) * (len(t.sp_class) + 1) // 2 for t in tables)
        assert len(got) == n


def test_tile_formats_follow_the_thresholds():
    """Format codes by tile max |element|: zero < tau_zero <= int8 < tau_i8
    <= int16 < tau_i16 <= FP32 < tau_f32 <= FP64; thresholds of 0 switch the
    integer formats off."""
    import numpy as np

    from jax_qc.integrals.store import (TILE_F32, TILE_F64, TILE_I16, TILE_I8, TILE_ZERO,
                                        tile_formats)

    tm = np.array([0.0, 1e-15, 1e-14, 5e-8, 1e-7, 2e-5, 3e-5, 5e-3, 1e-2, 1.0])
    f = tile_formats(tm, tau_f32=1e-2, tau_zero=1e-14, tau_i16=3e-5, tau_i8=1e-7)
    assert list(f) == [TILE_ZERO, TILE_ZERO, TILE_I8, TILE_I8, TILE_I16, TILE_I16, TILE_F32,
                       TILE_F32, TILE_F64, TILE_F64]
    f0 = tile_formats(tm, tau_f32=1e-2, tau_zero=1e-14)
    assert list(f0) == [TILE_ZERO, TILE_ZERO] + [TILE_F32] * 6 + [TILE_F64] * 2
