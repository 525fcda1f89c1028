"""GPU parity of the integral kernels against the CPU oracle.

Bars: Boys FP64 2e-14 relative vs mpmath; one-electron matrices and ERIs
1e-12 absolute vs the oracle's accurate-Boys restatement (the north-star
bar is 1e-10); Schwarz bounds 1e-12 relative.  The full-size (H2O)_8 store
is checked on a seeded sample of elements against the oracle's contracted
blocks."""

import json
import os

import numpy as np
import pytest
import torch

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from oracle import mmd

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "boys_mpmath.json")))


def _dev():
    return torch.device("cuda", 0)


def test_boys_fp64_matches_mpmath():
    from jax_qc.integrals.boys import boys_table

    t = np.array(GOLD["t"])
    F = np.array(GOLD["F"])
    got = boys_table(12, t).cpu().numpy()
    np.testing.assert_allclose(got, F, rtol=2e-14, atol=1e-300)


def test_boys_fp32_matches_mpmath():
    from jax_qc.integrals.boys import boys_table

    t = np.array(GOLD["t"])
    F = np.array(GOLD["F"])
    got = boys_table(8, t, dtype=torch.float32).double().cpu().numpy()
    np.testing.assert_allclose(got, F[:, :9], rtol=3e-6, atol=1e-37)


def test_boys_dense_grid_against_oracle():
    from jax_qc.integrals.boys import boys_table

    t = np.concatenate([np.linspace(0, 45, 4001), np.geomspace(1e-10, 1e3, 500)])
    got = boys_table(12, t).cpu().numpy()
    for n in (0, 3, 8, 12):
        np.testing.assert_allclose(got[:, n], mmd.boys(n, t), rtol=3e-14, atol=1e-300)


def test_boys_rejects_negative_n():
    from jax_qc.integrals.boys import boys_fn

    with pytest.raises(ValueError):
        boys_fn(-1, 1.0)


CASES = [("h2o-sto3g", M.h2o_tier2, "sto-3g"), ("h2o-631gss", M.h2o_tier2, "6-31g**"),
         ("water-et", M.water_monomer, "even-tempered")]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    _, mk, b = request.param
    mol = mk()
    bs = build_basis_set(mol, b)
    print(f"[case] {request.param[0]} nbf={bs.n_basis}", flush=True)
    return mol, bs, mmd.compute_integrals(mol, bs)


def test_one_electron_parity(case):
    from jax_qc.integrals import (compute_kinetic_matrix, compute_nuclear_matrix,
                                  compute_overlap_matrix)

    mol, bs, ref = case
    np.testing.assert_allclose(compute_overlap_matrix(bs).cpu().numpy(), ref["S"], atol=1e-12)
    np.testing.assert_allclose(compute_kinetic_matrix(bs).cpu().numpy(), ref["T"], atol=1e-12)
    np.testing.assert_allclose(compute_nuclear_matrix(bs, mol).cpu().numpy(), ref["V"], atol=1e-11)


@pytest.mark.parametrize("variant", [0, 1, 2], ids=["default", "per-thread", "per-wave"])
def test_dense_eri_parity(case, variant):
    from jax_qc.integrals import compute_eri_tensor
    from jax_qc.integrals import engine as eng

    mol, bs, ref = case
    eng.FORCE_VARIANT = variant
    try:
        got = compute_eri_tensor(bs).cpu().numpy()
        got32 = compute_eri_tensor(bs, tau32=np.inf).cpu().numpy()
    finally:
        eng.FORCE_VARIANT = None
    np.testing.assert_allclose(got, ref["ERI"], atol=1e-12, rtol=0)
    assert np.abs(got32 - ref["ERI"]).max() <= 5e-6 * max(1.0, np.abs(ref["ERI"]).max())


@pytest.mark.parametrize("variant", [1, 2], ids=["per-thread", "per-wave"])
def test_schwarz_parity(case, variant):
    from jax_qc.integrals import compute_shell_pair_bounds
    from jax_qc.integrals import engine as eng

    mol, bs, ref = case
    eng.FORCE_VARIANT = variant
    try:
        got = compute_shell_pair_bounds(bs)
    finally:
        eng.FORCE_VARIANT = None
    np.testing.assert_allclose(got, mmd.shell_pair_bounds(bs), rtol=1e-12, atol=1e-14)


def test_fp32_eri_error_is_bounded(case):
    from jax_qc.integrals import compute_eri_tensor

    mol, bs, ref = case
    # every quartet in FP32 (tau32 = inf): relative error of an FP32 sum
    got = compute_eri_tensor(bs, tau32=np.inf).cpu().numpy()
    err = np.abs(got - ref["ERI"])
    assert err.max() <= 5e-6 * max(1.0, np.abs(ref["ERI"]).max())


def test_store_fock_matches_oracle_einsum(case):
    from jax_qc.integrals.store import ERIStore
    from oracle import scf

    mol, bs, ref = case
    n = bs.n_basis
    rng = np.random.default_rng(5)
    A = rng.normal(size=(n, n))
    D = (A + A.T) / 2
    st = ERIStore(bs, tau=0.0)
    Dt = torch.from_numpy(D).to(_dev())
    Ht = torch.from_numpy(ref["H_core"]).to(_dev())
    F = st.fock_rhf(Ht, Dt).cpu().numpy()
    np.testing.assert_allclose(F, scf.build_fock_rhf(ref["H_core"], D, ref["ERI"]), atol=1e-11)
    np.testing.assert_allclose(st.coulomb(Dt).cpu().numpy(), scf.coulomb(D, ref["ERI"]), atol=1e-11)
    np.testing.assert_allclose(st.exchange(Dt).cpu().numpy(), scf.exchange(D, ref["ERI"]), atol=1e-11)


def test_store_uhf_fock_from_p_and_k_matches_oracle(case):
    """UHF Fock matrices from the J and K supermatrices and, after J is turned
    into P = J - K/2 in place (assemble_p_k, the energy + gradient setup), from
    P(Da + Db) -+ K(Da - Db)/2, both against the oracle's einsum build; J(D)
    then comes from P + K/2, and P equals the gathered assembly of P from J."""
    from jax_qc.integrals.store import ERIStore
    from oracle import scf

    mol, bs, ref = case
    n = bs.n_basis
    rng = np.random.default_rng(6)
    Da, Db = [(A + A.T) / 2 for A in rng.normal(size=(2, n, n))]
    want = scf.build_fock_uhf(ref["H_core"], Da, Db, ref["ERI"])
    st = ERIStore(bs, tau=0.0)
    dev = _dev()
    Ht = torch.from_numpy(ref["H_core"]).to(dev)
    Dat, Dbt = torch.from_numpy(Da).to(dev), torch.from_numpy(Db).to(dev)
    for form in ("J and K", "P and K"):
        if form == "P and K":
            Pg = st.ensure_P().clone()  # gathered from J
            st.P = None
            st.assemble_p_k()
            assert st.J is None and st.P is not None
            np.testing.assert_allclose(st.P.cpu().numpy(), Pg.cpu().numpy(), atol=1e-15, rtol=1e-15)
            del Pg
        Fa, Fb = st.fock_uhf(Ht, Dat, Dbt)
        np.testing.assert_allclose(Fa.cpu().numpy(), want[0], atol=1e-11, err_msg=form)
        np.testing.assert_allclose(Fb.cpu().numpy(), want[1], atol=1e-11, err_msg=form)
    np.testing.assert_allclose(st.coulomb(Dat).cpu().numpy(), scf.coulomb(Da, ref["ERI"]), atol=1e-11)


def test_assembly_tile_maxima_equal_tile_maxabs_of_the_result(case):
    """The per-tile max |element| the assembly kernels write beside P (kind 0),
    K (kind 1) and P = J - K/2 in place equals qc_sm_tile_maxabs over the
    assembled matrix bit for bit, so a compaction from them is the same
    compaction (directory and pool) as one that reads the matrix again."""
    from jax_qc import _native as nat
    from jax_qc.integrals.store import ERIStore

    mol, bs, ref = case
    st = ERIStore(bs, tau=0.0)

    def maxabs(M):
        t = st._tile_vector()
        nat.call("qc_sm_tile_maxabs", nat.ptr(M), st.nbf, nat.ptr(t), nat.stream_ptr())
        return t

    st.ensure_P()
    st.ensure_K()
    assert torch.equal(st._tmax["P"], maxabs(st.P))
    assert torch.equal(st._tmax["K"], maxabs(st.K))
    assert float(st._tmax["K"].max()) > 0.0
    st.P = None
    st.assemble_p_k()
    assert torch.equal(st._tmax["P"], maxabs(st.P))
    kw = dict(tau_f32=1e-2, tau_i16=3e-5, tau_i8=1e-7)
    for name, M in (("P", st.P), ("K", st.K)):
        a = st.compact(M, tmax=st._tmax[name], **kw)
        b = st.compact(M, **kw)
        assert torch.equal(a.fmt, b.fmt) and torch.equal(a.off, b.off), name
        assert torch.equal(a.pool, b.pool) and torch.equal(a.tmax, b.tmax), name


def test_dense_jk_kernel(case):
    from jax_qc.scf.fock import _dense_jk
    from oracle import scf

    mol, bs, ref = case
    n = bs.n_basis
    D = np.random.default_rng(1).normal(size=(n, n))
    D = D + D.T
    J, K = _dense_jk(torch.from_numpy(D).to(_dev()), torch.from_numpy(ref["ERI"]).to(_dev()))
    np.testing.assert_allclose(J.cpu().numpy(), scf.coulomb(D, ref["ERI"]), atol=1e-11)
    np.testing.assert_allclose(K.cpu().numpy(), scf.exchange(D, ref["ERI"]), atol=1e-11)


def test_water8_store_sampled_elements():
    """Full-size (H2O)_8 / even-tempered (240 functions): 300 seeded
    elements of the J supermatrix against the oracle's contracted blocks."""
    from jax_qc.integrals import engine as eng
    from jax_qc.integrals.store import ERIStore

    mol = M.water_cluster((2, 2, 2), seed=0)
    bs = build_basis_set(mol, "even-tempered")
    st = ERIStore(bs, tau=0.0)
    J = st.J.cpu().numpy()
    ao = np.cumsum([0] + [2 * s.angular_momentum + 1 for s in bs.shells])
    shell_of = np.searchsorted(ao, np.arange(bs.n_basis), side="right") - 1
    rng = np.random.default_rng(11)
    C = {2: np.array([[0, 3 ** .5, 0, 0, 0, 0], [0, 0, 0, 0, 3 ** .5, 0], [-.5, 0, 0, -.5, 0, 1],
                      [0, 0, 3 ** .5, 0, 0, 0], [3 ** .5 / 2, 0, 0, -3 ** .5 / 2, 0, 0]]),
         1: np.eye(3), 0: np.eye(1)}

    def pos(p, q):
        if p < q:
            p, q = q, p
        R, Cc = p // 64, q // 64
        return (R * (R + 1) // 2 + Cc) * 4096 + (p % 64) * 64 + q % 64

    worst = 0.0
    for _ in range(300):
        i, j, k, l = rng.integers(0, bs.n_basis, 4)
        si, sj, sk, sl = (int(shell_of[x]) for x in (i, j, k, l))
        blk = mmd.contracted_eri_block(bs, si, sj, sk, sl)
        ls = [bs.shells[s].angular_momentum for s in (si, sj, sk, sl)]
        sph = np.einsum("Ii,Jj,Kk,Ll,ijkl->IJKL", C[ls[0]], C[ls[1]], C[ls[2]], C[ls[3]], blk)
        ref = sph[i - ao[si], j - ao[sj], k - ao[sk], l - ao[sl]]
        p = max(i, j) * (max(i, j) + 1) // 2 + min(i, j)
        q = max(k, l) * (max(k, l) + 1) // 2 + min(k, l)
        worst = max(worst, abs(J[pos(p, q)] - ref))
    assert worst < 1e-12
