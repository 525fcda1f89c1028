"""Mixed precision: FP32-split ERIs and FP32 supermatrix tiles against the
FP64 path (bars: tile-exact when every tile is FP64; |dE| <= 1 uEh for the
default thresholds on a water cluster; FP32 storage error bounded by the
FP32 rounding of the stored values)."""

import numpy as np
import pytest
import torch

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from jax_qc.core.types import CalcConfig

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dimer():
    mol = M.water_cluster((2, 1, 1), seed=4)
    return mol, build_basis_set(mol, "even-tempered")


def test_compact_all_fp64_is_exact(dimer):
    from jax_qc.integrals.store import ERIStore

    mol, bs = dimer
    st = ERIStore(bs, tau=1e-12)
    P = st.ensure_P()
    C = st.compact(P, tau_f32=0.0, tau_zero=0.0)
    assert C.n_f32 == 0 and C.n_zero == 0
    n = bs.n_basis
    D = torch.randn(n, n, dtype=torch.float64, device="cuda")
    D = D + D.T
    a = st.contract(P, D)
    b = st.contract(C, D)
    assert torch.equal(a, b)


def test_compact_fp32_error_bounded(dimer):
    from jax_qc.integrals.store import ERIStore

    mol, bs = dimer
    st = ERIStore(bs, tau=1e-12)
    P = st.ensure_P()
    C = st.compact(P, tau_f32=np.inf, tau_zero=0.0)
    assert C.n_f64 == 0
    n = bs.n_basis
    D = torch.randn(n, n, dtype=torch.float64, device="cuda")
    D = D + D.T
    a = st.contract(P, D)
    b = st.contract(C, D)
    # each element rounded to FP32 (rel 6e-8); G_ij sums npair of them
    bound = 6e-8 * st.contract(st.compact(P.abs(), tau_f32=0.0), D.abs()).max()
    assert (a - b).abs().max() <= bound


@pytest.mark.parametrize("fmt", ["i16", "i8", "all"])
def test_compact_integer_tiles_error_bounded(dimer, fmt):
    """int16 / int8 tiles: every element is off by at most max|tile| / 65534
    (int16) or / 254 (int8), so y = M x is off by at most the contraction of
    |x| with a supermatrix of those per-tile bounds."""
    from jax_qc.integrals.store import ERIStore, TILE_F32, TILE_F64, TILE_I16, TILE_I8

    mol, bs = dimer
    st = ERIStore(bs, tau=1e-12)
    P = st.ensure_P()
    if fmt == "all":
        # all four formats: thresholds at the quartiles of the tile maxima
        q = np.quantile(st.compact(P, tau_f32=0.0, tau_zero=0.0).tmax.cpu().numpy(),
                        [0.25, 0.5, 0.75])
        kw = dict(tau_i8=q[0], tau_i16=q[1], tau_f32=q[2])
    else:
        kw = {"i16": dict(tau_f32=np.inf, tau_i16=np.inf),
              "i8": dict(tau_f32=np.inf, tau_i16=np.inf, tau_i8=np.inf)}[fmt]
    C = st.compact(P, tau_zero=0.0, **kw)
    f = C.fmt.cpu().numpy()
    if fmt == "i16":
        assert C.n_i16 == len(f) and C.n_i8 == 0
    elif fmt == "i8":
        assert C.n_i8 == len(f)
    else:
        assert C.n_i16 > 0 and C.n_i8 > 0 and C.n_f32 > 0 and C.n_f64 > 0
    n = bs.n_basis
    g = torch.Generator(device="cpu").manual_seed(3)
    D = torch.randn(n, n, dtype=torch.float64, generator=g).cuda()
    D = D + D.T
    a = st.contract(P, D)
    b = st.contract(C, D)
    tm = C.tmax.double()
    fd = C.fmt.long()
    per = torch.zeros_like(tm)
    per[fd == TILE_I16] = tm[fd == TILE_I16] / 65534.0
    per[fd == TILE_I8] = tm[fd == TILE_I8] / 254.0
    per[fd == TILE_F32] = tm[fd == TILE_F32] * 6e-8
    per[fd == TILE_F64] = tm[fd == TILE_F64] * 1e-14  # summation order only
    B = per.repeat_interleave(64 * 64)
    assert B.numel() == P.numel()
    bound = st.contract(B, D.abs()).abs() * (1 + 1e-6) + 1e-15
    err = (a - b).abs()
    assert bool((err <= bound).all()), float((err - bound).max())
    assert float(err.max()) > 0.0  # the quantisation is in effect


def test_mixed_scf_energy_within_1uEh(dimer):
    import jax_qc

    mol, bs = dimer
    E = {}
    for mode in ("fp64", "mixed"):
        ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-12,
                                        tau32=1e-4 if mode == "mixed" else None)
        if mode == "mixed":
            ints.ERI.use_compact(tau_f32=1e-4)
            assert ints.ERI.plan.n32 > 0
        r = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf"))
        assert r.converged
        E[mode] = r.energy
    assert abs(E["mixed"] - E["fp64"]) <= 1e-6


def test_autotune_picks_a_valid_variant(dimer):
    from jax_qc import autotune
    from jax_qc.integrals import engine as eng

    mol, bs = dimer
    pt = eng.PairTable(bs)
    plan = eng.plan_quartets(pt, 1e-12)
    saved = dict(eng.ERI_VARIANT)
    try:
        t = autotune.autotune_eri(pt, plan, repeats=1)
        assert t and all(set(v) == {1, 2} for v in t.values())
        assert all(eng.ERI_VARIANT[(a, b)] in (1, 2) for (a, b, _) in t)
    finally:
        eng.ERI_VARIANT.clear()
        eng.ERI_VARIANT.update(saved)


@pytest.mark.parametrize("nrhs", [1, 2, 3])
def test_symv_two_wave_items_match_one_wave(dimer, nrhs):
    """The tile pass with two waves per work item (each half of a tile's
    rows, column sums combined through LDS) gives the one-wave result to
    rounding, on the plain and the compacted (screened) supermatrix."""
    from jax_qc import _native as nat
    from jax_qc.integrals.store import ERIStore

    mol, bs = dimer
    st = ERIStore(bs, tau=1e-12)
    P = st.ensure_P()
    C = st.compact(P, tau_f32=1e-6)
    n = bs.n_basis
    g = torch.Generator(device="cpu").manual_seed(nrhs)
    A = torch.randn(nrhs, n, n, generator=g, dtype=torch.float64)
    D = (A + A.transpose(1, 2)).cuda()
    out = {}
    try:
        for ws in (1, 2):
            nat.call("qc_symv_split", nrhs, ws)
            out[ws] = (st.contract(P, D), st.contract(C, D), st.contract(C, D, tau_skip=1e-9))
    finally:
        nat.call("qc_symv_split", nrhs, 1 if nrhs == 1 else 2)
    for a, b in zip(out[1], out[2]):
        scale = float(a.abs().max())
        assert float((a - b).abs().max()) <= 1e-13 * scale


@pytest.mark.parametrize("method", ["rhf", "uhf", "uhf-pk"])
def test_incremental_fock_builds_and_their_graphs(method):
    """Incremental Fock builds (difference densities, tiles screened at
    1e-13) land within 1e-8 Eh of full builds on the mixed store, the
    converged result rests on a full build, and replaying the builds from
    HIP graphs changes nothing (1e-10 Eh, same iteration count)."""
    import jax_qc
    from jax_qc.integrals import store as stmod

    spin = 2 if method.startswith("uhf") else 0
    mol = M.water_cluster((2, 1, 1), seed=1, spin=spin)
    bs = build_basis_set(mol, "even-tempered")
    cfg = CalcConfig(method=method[:3], scf_conv=1e-10, max_scf_iter=200)

    def run(inc, graphs):
        ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-12, tau32=1e-4)
        if method == "uhf":  # J and K pools
            ints.ERI.ensure_K()
        elif method == "uhf-pk":  # P = J - K/2 and K pools (the energy + gradient setup)
            ints.ERI.assemble_p_k()
        else:
            ints.ERI.ensure_P()
            ints.ERI.release_J()
        ints.ERI.use_compact(tau_f32=1e-2)
        if inc:
            ints.ERI.set_incremental(1e-13)
        stmod.USE_FOCK_GRAPHS = graphs
        try:
            r = jax_qc.run_scf(mol, ints, cfg)
        finally:
            stmod.USE_FOCK_GRAPHS = True
        assert r.converged
        return r, ints.ERI.inc

    full, _ = run(False, False)
    eager, inc_e = run(True, False)
    graph, inc_g = run(True, True)
    print(method, "full", full.energy, "inc", eager.energy, "graph", graph.energy,
          "builds", inc_g.n_full, inc_g.n_inc, flush=True)
    assert inc_e.n_inc > 0 and not inc_e.last_incremental
    assert abs(eager.energy - full.energy) < 1e-8
    assert abs(graph.energy - eager.energy) < 1e-10
    assert graph.n_iterations == eager.n_iterations
    assert set(inc_g.graphs) == {"full", "inc"} or inc_g.n_full < 2
