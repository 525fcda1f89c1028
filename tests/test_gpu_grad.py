"""Analytic gradient (hand-written backward kernels) against
* the oracle's restatement of the reference's rhf_gradient (FD of full
  integral rebuilds at fixed D, W; jax_qc/grad/rhf_grad.py:106-211),
* full-SCF central differences at h = 1e-4 Bohr (numerical_grad.py:51-94),
* invariants (sum over atoms = 0, H2 antisymmetry; test_gradient.py:119-133).
Bars: 1e-7 Eh/Bohr against the FD-of-integrals restatement (its O(h^2) error
is ~1e-9); 1e-6 against full-SCF differences."""

import numpy as np
import pytest
import torch

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from jax_qc.core.types import CalcConfig, make_molecule
from oracle import mmd, scf

pytestmark = pytest.mark.gpu


def _displace(mol, c):
    return make_molecule(mol.elements, c, mol.atomic_numbers, mol.charge, mol.spin)


def _gpu_grad(mol, basis, method="rhf", tau=0.0):
    import jax_qc
    from jax_qc.grad import rhf_gradient

    bs = build_basis_set(mol, basis)
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=tau)
    res = jax_qc.run_scf(mol, ints, CalcConfig(method=method, scf_conv=1e-12))
    assert res.converged
    return rhf_gradient(mol, res, basis, basis=bs).cpu().numpy(), res


CASES = [("h2-sto3g", M.h2, "sto-3g"), ("h2o-631gss", M.h2o_tier2, "6-31g**"),
         ("water-et", M.water_monomer, "even-tempered")]


@pytest.mark.parametrize("name,mk,basis", CASES, ids=[c[0] for c in CASES])
def test_rhf_gradient_matches_oracle_fd_of_integrals(name, mk, basis):
    mol = mk()
    print(f"[case] {name}", flush=True)
    g, _ = _gpu_grad(mol, basis)
    bf = lambda m: build_basis_set(m, basis)
    ref = scf.run_rhf(mol, mmd.compute_integrals(mol, bf(mol)), conv=1e-12)
    go = scf.rhf_gradient(mol, ref, bf, _displace, step=1e-4)
    np.testing.assert_allclose(g, go, atol=1e-7)
    np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-9)


def test_h2_gradient_antisymmetric_z_only():
    g, _ = _gpu_grad(M.h2(), "sto-3g")
    assert abs(g[0, 2] + g[1, 2]) < 1e-10
    assert np.abs(g[:, :2]).max() < 1e-12


def test_rhf_gradient_matches_full_scf_fd():
    from jax_qc.grad import numerical_gradient

    mol = M.h2o_tier2()
    g, _ = _gpu_grad(mol, "6-31g**")
    gn = numerical_gradient(mol, "6-31g**", "rhf", 1e-4, scf_conv=1e-12).numpy()
    np.testing.assert_allclose(g, gn, atol=1e-6)


def test_uhf_gradient_matches_full_scf_fd():
    from jax_qc.grad import numerical_gradient

    mol = M.oh_radical()
    g, res = _gpu_grad(mol, "sto-3g", method="uhf")
    gn = numerical_gradient(mol, "sto-3g", "uhf", 1e-4, scf_conv=1e-12).numpy()
    np.testing.assert_allclose(g, gn, atol=1e-6)


def test_autograd_backward_equals_analytic():
    from jax_qc.grad import scf_energy

    mol = M.h2o_tier2()
    x = torch.tensor(mol.coords, dtype=torch.float64, requires_grad=True)
    E = scf_energy(x, mol, "6-31g", CalcConfig(method="rhf", scf_conv=1e-12), tau=0.0)
    E.backward()
    g, res = _gpu_grad(mol, "6-31g")
    assert float(E) == pytest.approx(res.energy, abs=1e-10)
    np.testing.assert_allclose(x.grad.numpy(), g, atol=1e-10)


def test_water4_gradient_all_36_coordinates_vs_fd():
    """(H2O)_4 / even-tempered (seed 1, 120 functions): analytic gradient on
    all 36 coordinates against full-SCF central differences, h = 1e-4."""
    from jax_qc.grad import numerical_gradient

    mol = M.water_cluster((2, 2, 1), seed=1)
    g, _ = _gpu_grad(mol, "even-tempered", tau=1e-14)
    gn = numerical_gradient(mol, "even-tempered", "rhf", 1e-4, scf_conv=1e-12, tau=1e-14).numpy()
    print("max |analytic - FD| =", np.abs(g - gn).max(), flush=True)
    np.testing.assert_allclose(g, gn, atol=1e-6)
    np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-8)


def test_kernel_forms_fp32_split_and_density_screening():
    """The derivative-ERI kernel forms (one thread / one wave / one lane per
    primitive quartet / one wave with the G contraction on the matrix cores /
    one lane per primitive quartet with G contracted first) agree to rounding; the FP32 split (Schwarz bound < 1e-4; either
    form's FP32 sums) and
    density-weighted screening (1e-11) stay within 1e-7 Eh/Bohr of FP64 on
    (H2O)_2 even-tempered (seed 1)."""
    import jax_qc
    from jax_qc.grad import analytic as an

    mol = M.water_cluster((2, 1, 1), seed=1)
    bs = build_basis_set(mol, "even-tempered")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-14)
    res = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf", scf_conv=1e-12))
    pt = ints.ERI.pt
    g = {}
    try:
        for ov in (1, 2, 3, 4, 5):
            an.GRAD_VARIANT_OVERRIDE = ov
            g[ov] = an.scf_gradient(mol, bs, res, 1e-14, pair_table=pt).cpu().numpy()
    finally:
        an.GRAD_VARIANT_OVERRIDE = None
    for ov in (1, 3, 4, 5):
        np.testing.assert_allclose(g[ov], g[2], atol=1e-11)
    g32 = an.scf_gradient(mol, bs, res, 1e-14, pair_table=pt, tau32=1e-4).cpu().numpy()
    print("max |fp32 split - fp64| =", np.abs(g32 - g[2]).max(), flush=True)
    np.testing.assert_allclose(g32, g[2], atol=1e-7)
    # the FP32 sums of each kernel form on their own
    try:
        for ov in (1, 2, 3, 4, 5):
            an.GRAD_VARIANT_OVERRIDE = ov
            g32f = an.scf_gradient(mol, bs, res, 1e-14, pair_table=pt, tau32=1e-4).cpu().numpy()
            print("form", ov, "max |fp32 split - fp64| =", np.abs(g32f - g[2]).max(), flush=True)
            np.testing.assert_allclose(g32f, g[2], atol=1e-7)
    finally:
        an.GRAD_VARIANT_OVERRIDE = None
    gd = an.scf_gradient(mol, bs, res, 1e-14, pair_table=pt, tau_density=1e-11).cpu().numpy()
    np.testing.assert_allclose(gd, g[2], atol=1e-7)


def test_thread_kernels_lds_staged_adds_equal_global_adds():
    """The thread-per-quartet kernels' per-work-group LDS sums (natom <= 256)
    give the gradient that direct global atomic adds give (natom = 0)."""
    import jax_qc
    from jax_qc.grad import analytic as an

    mol = M.water_cluster((2, 1, 1), seed=1)
    bs = build_basis_set(mol, "even-tempered")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-14)
    res = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf", scf_conv=1e-12))
    pt = ints.ERI.pt
    g = {}
    try:
        an.GRAD_VARIANT_OVERRIDE = 1
        for staged in (True, False):
            an.GRAD_STAGE_IN_LDS = staged
            g[staged] = an.scf_gradient(mol, bs, res, 1e-14, pair_table=pt).cpu().numpy()
    finally:
        an.GRAD_VARIANT_OVERRIDE = None
        an.GRAD_STAGE_IN_LDS = True
    np.testing.assert_allclose(g[True], g[False], atol=1e-12)


@pytest.mark.parametrize("method", ["rhf", "uhf"])
def test_density_prescreen_leaves_gradient_unchanged(method):
    """The launch-side density prescreen (shell-block density maxima) drops
    only quartets the kernels' own screen skips: the gradient with it equals
    the gradient without it, in FP64 and with the FP32 split, RHF and UHF
    ((H2O)_2 even-tempered with d shells)."""
    import jax_qc
    from jax_qc.grad import analytic as an

    spin = 2 if method == "uhf" else 0
    mol = M.water_cluster((2, 1, 1), seed=1, spin=spin)
    bs = build_basis_set(mol, "even-tempered")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-14)
    res = jax_qc.run_scf(mol, ints, CalcConfig(method=method, scf_conv=1e-12, max_scf_iter=200))
    pt = ints.ERI.pt
    for tau32 in (None, 1e-2):
        g = {}
        try:
            for pre in (True, False):
                an.PRESCREEN = pre
                g[pre] = an.scf_gradient(mol, bs, res, 1e-12, pair_table=pt, tau_density=1e-9,
                                         tau32=tau32).cpu().numpy()
        finally:
            an.PRESCREEN = True
        np.testing.assert_allclose(g[True], g[False], atol=1e-13, rtol=0)


@pytest.mark.parametrize("nk", [1, 2])
def test_prescreen_kernel_keeps_exactly_the_bounded_quartets(nk):
    """qc_grad2e_prescreen against the bound evaluated in NumPy for every
    quartet of every launch ((H2O)_2 even-tempered, random shell-block
    maxima spanning 1e-8..1, one or two exchange densities): the kept ket
    lists, per launch and per bra in list order, are those with
    bound * Q_ab Q_cd (1 + 1e-6) >= tau, apart from quartets within 1e-9
    (relative) of the threshold, which may go either way."""
    import jax_qc
    from jax_qc.grad import analytic as an
    from jax_qc.integrals import engine as eng

    mol = M.water_cluster((2, 1, 1), seed=1)
    bs = build_basis_set(mol, "even-tempered")
    pt = eng.PairTable(bs, torch.device("cuda", 0))
    plan = eng.plan_quartets(pt, 1e-12, 1e-4)
    ns = pt.nshell
    rng = np.random.default_rng(7)

    def sym(x):
        return np.maximum(x, x.T)

    Mj = sym(10.0 ** rng.uniform(-8, 0, (ns, ns)))
    Mk = [sym(10.0 ** rng.uniform(-8, 0, (ns, ns))) for _ in range(nk)]
    cx, tau = 0.5, 1e-9
    Q = pt.Q.astype(np.float64)
    shells = pt.sp_shells.astype(np.int64)
    dev = torch.device("cuda", 0)
    got = an.prescreen_launches(plan.launches, pt, torch.from_numpy(shells).to(dev),
                                torch.from_numpy(Mj).to(dev), [torch.from_numpy(m).to(dev) for m in Mk],
                                cx, torch.from_numpy(Q).to(dev), tau)
    got = {(L.cab, L.ccd, L.fp32): L for L in got}
    n_checked = n_dropped = 0
    for L in plan.launches:
        bra = L.bra.cpu().numpy()
        ket = L.ket.cpu().numpy()
        offs = L.offsets.cpu().numpy()
        ks = L.kstart.cpu().numpy()
        G = got.get((L.cab, L.ccd, L.fp32))
        if G is not None:
            go, gk = G.offsets.cpu().numpy(), G.ket.cpu().numpy()
        for i in range(L.nbra):
            kp = ket[ks[i]:ks[i] + offs[i + 1] - offs[i]]
            a, b = shells[bra[i]]
            c, d = shells[kp, 0], shells[kp, 1]
            bound = Mj[a, b] * Mj[c, d]
            for m in Mk:
                bound = bound + cx * (m[a, c] * m[b, d] + m[a, d] * m[b, c])
            val = bound * Q[bra[i]] * Q[kp] * (1.0 + 1e-6)
            sure = np.abs(val / tau - 1.0) > 1e-9
            want = kp[(val >= tau) & sure]
            if G is None:
                assert not np.any((val >= tau) & sure), (L.cab, L.ccd, i)
                continue
            kept = gk[go[i]:go[i + 1]]
            assert np.all(np.isin(want, kept)), (L.cab, L.ccd, L.fp32, i)
            assert not np.any(np.isin(kp[(val < tau) & sure], kept)), (L.cab, L.ccd, L.fp32, i)
            order = {k: j for j, k in enumerate(kp)}
            assert all(order[x] < order[y] for x, y in zip(kept[:-1], kept[1:]))  # list order
            n_checked += len(kp)
            n_dropped += len(kp) - len(kept)
    assert n_checked > 10000 and 0 < n_dropped < n_checked
