"""The benchmark configurations on the GPU, each pinned to the oracle or to
properties fixed by the mathematics:

* C1: the H2O / STO-3G geometry of the benchmark (O at the origin, H at
  (0, +-0.7572, 0.5865) A) against the oracle's RHF (rhf.py:52-169).
* C2: (H2O)_8 mixed path (FP32-split ERIs, FP32 tiles) within 1 uEh of the
  FP64 path.
* C3: (H2O)_4 analytic gradient against the oracle's restatement of the
  reference's gradient (FD of integrals at fixed D, W; rhf_grad.py:106-211)
  on all 36 coordinates, from a committed fixture
  (tests/golden/make_scf_golden.py).
* C4: all 256 batched molecules, graph replays equal eager builds, mixed
  within 1 uEh of FP64; six batched molecules against oracle energies
  (fixture).
* C5: (H2O)_16 UHF triplet, mixed vs FP64 energy and gradient (1 uEh, 1e-6
  Eh/Bohr), translational sum rule, energy invariant under a rigid motion.
* UHF with d shells: (H2O)_2 triplet energy against the oracle's UHF
  (uhf.py:187-287, fixture) and the UHF gradient (the d-class kernels with
  two exchange densities) against full-SCF central differences
  (numerical_grad.py:51-94) for both derivative-kernel forms and the FP32
  split.
"""

import json
import os

import numpy as np
import pytest
import torch

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from jax_qc.core.types import CalcConfig

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def _same_molecule(mol, rec):
    assert list(mol.elements) == rec["elements"]
    np.testing.assert_allclose(np.asarray(mol.coords), np.asarray(rec["coords_bohr"]), atol=1e-12)


def test_c1_water_monomer_sto3g_matches_oracle():
    import jax_qc
    from oracle import mmd, scf

    print("[case] C1 H2O/STO-3G", flush=True)
    mol = M.water_monomer()
    bs = build_basis_set(mol, "sto-3g")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=0.0)
    r = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf", scf_conv=1e-10, diis_space=8))
    ref = scf.run_rhf(mol, mmd.compute_integrals(mol, bs), conv=1e-10)
    assert r.converged and ref["converged"]
    assert abs(r.energy - ref["energy"]) < 1e-8


def test_c2_water8_mixed_within_a_microhartree():
    import jax_qc

    print("[case] C2 (H2O)_8 mixed vs FP64", flush=True)
    mol = M.water_cluster((2, 2, 2), seed=0)
    bs = build_basis_set(mol, "even-tempered")
    E = {}
    for mode in ("fp64", "mixed"):
        ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-12,
                                        tau32=1e-4 if mode == "mixed" else None)
        ints.ERI.release_J()
        if mode == "mixed":
            ints.ERI.use_compact(tau_f32=1e-2)
            assert ints.ERI.plan.n32 > 0 and ints.ERI.P.n_f32 > 0
        r = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf", scf_conv=1e-10))
        assert r.converged
        E[mode] = r.energy
        del ints
        torch.cuda.empty_cache()
    print("dE", abs(E["mixed"] - E["fp64"]), flush=True)
    assert abs(E["mixed"] - E["fp64"]) <= 1e-6


def test_c3_water4_gradient_matches_oracle_fixture():
    import jax_qc
    from jax_qc.grad import rhf_gradient

    print("[case] C3 (H2O)_4 gradient vs oracle FD-of-integrals fixture", flush=True)
    gold = _golden("water4_rhf_grad.json")
    mol = M.water_cluster((2, 2, 1), seed=1)
    _same_molecule(mol, gold["molecule"])
    bs = build_basis_set(mol, "even-tempered")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-14)
    res = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf", scf_conv=1e-12))
    assert res.converged
    assert abs(res.energy - gold["energy"]) < 1e-8
    g = rhf_gradient(mol, res, "even-tempered", basis=bs).cpu().numpy()
    go = np.asarray(gold["gradient"])
    print("max |analytic - oracle FD of integrals| =", np.abs(g - go).max(), flush=True)
    np.testing.assert_allclose(g, go, atol=1e-7)


def test_c4_batch6_energies_match_oracle_fixture():
    from jax_qc.batch import BatchedRHF

    print("[case] C4 six batched molecules vs oracle energies", flush=True)
    gold = _golden("batch6_rhf.json")
    mols = M.random_molecules(6, seed=0, max_atoms=6)
    for m, rec in zip(mols, gold["molecules"]):
        _same_molecule(m, rec["molecule"])
    r = BatchedRHF(mols, tau=1e-12).run(conv=1e-10)
    for i, rec in enumerate(gold["molecules"]):
        assert rec["converged"] and r.converged[i]
        assert abs(r.energies[i] - rec["energy"]) < 1e-8, (i, r.energies[i], rec["energy"])


def test_c4_batch256_graph_replay_and_mixed_precision():
    from jax_qc.batch import BatchedRHF

    print("[case] C4 256 molecules: graph replays, mixed vs FP64", flush=True)
    mols = M.random_molecules(256, seed=0)
    res = {}
    for mode in ("fp64", "mixed"):
        kw = dict(tau=1e-12) if mode == "fp64" else dict(tau=1e-12, tau32=1e-4, tile_f32=1e-2)
        b = BatchedRHF(mols, **kw)
        g = torch.Generator(device="cpu").manual_seed(3)
        for rep in range(3):
            D = torch.zeros(b.nmol, b.nmax, b.nmax, dtype=torch.float64)
            for i in range(b.nmol):
                k = int(b.nbf_m[i])
                A = torch.randn(k, k, generator=g, dtype=torch.float64)
                D[i, :k, :k] = 0.05 * (A + A.T)
            D = D.cuda()
            Fg = b.fock(D, use_graph=True)
            torch.cuda.synchronize()
            _ = (D @ D).sum().item()  # eager work between replays
            Fe = b.fock(D, use_graph=False)
            assert torch.equal(Fg, Fe), rep
        r = b.run(conv=1e-10)
        res[mode] = r
        print(mode, "converged", int(r.converged.sum()), "of", b.nmol, flush=True)
        del b
        torch.cuda.empty_cache()
    both = res["fp64"].converged & res["mixed"].converged
    assert both.sum() >= 240
    dE = np.abs(res["fp64"].energies[both] - res["mixed"].energies[both]).max()
    print("max |dE| mixed vs fp64 over", int(both.sum()), "molecules:", dE, flush=True)
    assert dE <= 1e-6


def _uhf_dimer():
    return M.water_cluster((2, 1, 1), seed=1, spin=2)


def test_uhf_water2_triplet_energy_matches_oracle_fixture():
    import jax_qc

    print("[case] (H2O)_2 UHF triplet vs oracle", flush=True)
    gold = _golden("water2_uhf_triplet.json")
    mol = _uhf_dimer()
    _same_molecule(mol, gold["molecule"])
    bs = build_basis_set(mol, "even-tempered")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-14)
    r = jax_qc.run_scf(mol, ints, CalcConfig(method="uhf", scf_conv=1e-12, max_scf_iter=200))
    assert r.converged
    assert abs(r.energy - gold["energy"]) < 1e-8, (r.energy, gold["energy"])


def test_uhf_water2_triplet_gradient_vs_full_scf_fd():
    """UHF gradient with d shells (both exchange densities, cx = 1/2) of
    every kernel form against full-SCF central differences (h = 1e-4)."""
    import jax_qc
    from jax_qc.grad import analytic as an
    from jax_qc.grad import numerical_gradient

    print("[case] (H2O)_2 UHF triplet gradient vs FD", flush=True)
    mol = _uhf_dimer()
    bs = build_basis_set(mol, "even-tempered")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-14)
    res = jax_qc.run_scf(mol, ints, CalcConfig(method="uhf", scf_conv=1e-12, max_scf_iter=200))
    assert res.converged
    gn = numerical_gradient(mol, "even-tempered", "uhf", 1e-4, scf_conv=1e-12, tau=1e-14).numpy()
    pt = ints.ERI.pt
    forms = {}
    try:
        for ov in (None, 1, 2, 3, 4, 5):
            an.GRAD_VARIANT_OVERRIDE = ov
            forms[ov] = an.scf_gradient(mol, bs, res, 1e-14, pair_table=pt).cpu().numpy()
            forms[(ov, "fp32")] = an.scf_gradient(mol, bs, res, 1e-14, pair_table=pt,
                                                  tau32=1e-4).cpu().numpy()
    finally:
        an.GRAD_VARIANT_OVERRIDE = None
    for k, g in forms.items():
        err = np.abs(g - gn).max()
        print("form", k, "max |analytic - FD| =", err, flush=True)
        assert err < 1e-6, k
        np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-9)
    np.testing.assert_allclose(forms[1], forms[2], atol=1e-11)
    np.testing.assert_allclose(forms[3], forms[2], atol=1e-11)
    np.testing.assert_allclose(forms[4], forms[2], atol=1e-11)
    np.testing.assert_allclose(forms[5], forms[2], atol=1e-11)


def _rigid_motion(coords, seed=7):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    R = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                  [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                  [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])
    return coords @ R.T + np.array([0.7, -1.3, 2.1])


def test_c5_water16_uhf_mixed_vs_fp64():
    """(H2O)_16 UHF triplet, 480 functions, in the benchmark's two modes
    (jax_qc.precision): the mixed path's energy and gradient against the FP64
    path (1 uEh, 1e-6 Eh/Bohr); the gradient sums
    to zero over atoms; the FP64 energy is unchanged by a rigid rotation +
    translation of the cluster."""
    from jax_qc.geomopt.optimizer import energy_and_gradient
    from jax_qc.grad.numerical_grad import scf_energy_at

    print("[case] C5 (H2O)_16 UHF triplet mixed vs FP64", flush=True)
    mol = M.water_cluster((4, 2, 2), seed=2, spin=2)
    cfg = CalcConfig(method="uhf", basis="even-tempered", scf_conv=1e-10, max_scf_iter=200)
    c0 = np.asarray(mol.coords, dtype=np.float64)
    # the benchmark's own option sets (FP32-split ERIs, FP32 tiles, primitive-pair
    # screening, incremental Fock builds, screened FP32-split gradient)
    from jax_qc.precision import MIXED as mixed, STRICT as strict
    E64, g64, _, r64 = energy_and_gradient(c0, mol, "even-tempered", cfg, **strict)
    Em, gm, _, rm = energy_and_gradient(c0, mol, "even-tempered", cfg, **mixed)
    assert r64.converged and rm.converged
    dE, dg = abs(Em - E64), np.abs(gm - g64).max()
    print(f"dE {dE:.3e}  max dgrad {dg:.3e}  |g|max {np.abs(g64).max():.3e}", flush=True)
    assert dE <= 1e-6 and dg <= 1e-6
    np.testing.assert_allclose(g64.sum(axis=0), 0.0, atol=1e-8)
    np.testing.assert_allclose(gm.sum(axis=0), 0.0, atol=1e-8)
    Erot = scf_energy_at(mol, _rigid_motion(c0), "even-tempered", cfg, tau=1e-12)
    print(f"rigid-motion dE {abs(Erot - E64):.3e}", flush=True)
    assert abs(Erot - E64) < 1e-8
