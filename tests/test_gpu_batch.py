"""Batched RHF (many molecules per launch) against the single-molecule path
(run_scf on an ERIStore, the reference's RHF step order, rhf.py:52-169) and
against per-molecule Fock builds; graph replays checked one by one with eager
work in between."""

import numpy as np
import pytest
import torch

import jax_qc
from jax_qc import molecules as M
from jax_qc.batch import BatchedRHF
from jax_qc.core.types import CalcConfig

pytestmark = pytest.mark.gpu


def _mols(n=6, seed=0):
    return M.random_molecules(n, seed=seed, max_atoms=6)


def test_batched_fock_graph_matches_per_molecule_store():
    mols = _mols()
    print("[case] batched fock, 6 molecules", flush=True)
    b = BatchedRHF(mols, tau=1e-12)
    g = torch.Generator(device="cpu").manual_seed(1)
    refs = []
    for i, (m, bs) in enumerate(zip(mols, b.bases)):
        ints = jax_qc.compute_integrals(m, bs, eri="store", tau=1e-12)
        refs.append(ints)
    for rep in range(3):
        D = torch.zeros(b.nmol, b.nmax, b.nmax, dtype=torch.float64)
        for i in range(b.nmol):
            k = int(b.nbf_m[i])
            A = torch.randn(k, k, generator=g, dtype=torch.float64)
            D[i, :k, :k] = 0.1 * (A + A.T)
        D = D.cuda()
        F_graph = b.fock(D, use_graph=True)
        torch.cuda.synchronize()
        _ = (D @ D).sum().item()  # eager work between replays
        F_eager = b.fock(D, use_graph=False)
        for i in range(b.nmol):
            k = int(b.nbf_m[i])
            ref = refs[i].ERI.fock_rhf(refs[i].H_core, D[i, :k, :k].contiguous())
            np.testing.assert_allclose(F_graph[i, :k, :k].cpu().numpy(), ref.cpu().numpy(), atol=1e-11)
            np.testing.assert_allclose(F_eager[i, :k, :k].cpu().numpy(), ref.cpu().numpy(), atol=1e-11)
            pad = F_graph[i, k:, k:].cpu().numpy()
            if pad.size:
                np.testing.assert_allclose(pad, 1e4 * np.eye(pad.shape[0]))


def test_batched_rhf_energies_match_single_molecule_scf():
    mols = _mols()
    print("[case] batched rhf, 6 molecules", flush=True)
    b = BatchedRHF(mols, tau=1e-12)
    r = b.run(conv=1e-10)
    assert r.converged.all()
    for i, (m, bs) in enumerate(zip(mols, b.bases)):
        ints = jax_qc.compute_integrals(m, bs, eri="store", tau=1e-12)
        ref = jax_qc.run_scf(m, ints, CalcConfig(method="rhf", scf_conv=1e-10))
        assert ref.converged
        assert abs(r.energies[i] - ref.energy) < 1e-8, (i, r.energies[i], ref.energy)
        # same step rules, so the same iteration count, up to one step: the
        # batched path sums its Fock builds, DIIS products and padded
        # eigensolves in another order, and where |dE| lands next to the
        # 1e-10 threshold those rounding differences decide the last step
        assert abs(int(r.n_iterations[i]) - int(ref.n_iterations)) <= 1, (i, r.n_iterations[i],
                                                                         ref.n_iterations)


def test_batched_mixed_precision_within_a_microhartree():
    mols = M.random_molecules(8, seed=5)  # 2-12 atoms
    print("[case] batched mixed vs fp64, 8 molecules", flush=True)
    e64 = BatchedRHF(mols, tau=1e-12).run(conv=1e-10)
    bm = BatchedRHF(mols, tau=1e-12, tau32=1e-4, tile_f32=1e-2)
    em = bm.run(conv=1e-10)
    print("tiles", bm.tile_counts, "fp32 quartets", bm.plan.n32, flush=True)
    assert e64.converged.all() and em.converged.all()
    assert np.abs(em.energies - e64.energies).max() < 1e-6
    assert bm.plan.n32 > 0
