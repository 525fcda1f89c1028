"""GPU SCF parity: energies through the device integrals and Fock builds
against the oracle's restated SCF and the reference's hard-coded PySCF pins
(1e-8 Eh against the oracle)."""

import numpy as np
import pytest
import torch

from jax_qc import molecules as M
from jax_qc.basis import build_basis_set
from jax_qc.core.types import CalcConfig
from oracle import mmd, scf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("basis,pin", [("sto-3g", -73.2144728363), ("6-31g**", -74.5428909110)])
@pytest.mark.parametrize("eri", ["dense", "store"])
def test_rhf_h2o(basis, pin, eri):
    import jax_qc

    mol = M.h2o_tier2()
    bs = build_basis_set(mol, basis)
    ints = jax_qc.compute_integrals(mol, bs, eri=eri, tau=0.0)
    r = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf"))
    ref = scf.run_rhf(mol, mmd.compute_integrals(mol, bs))
    assert r.converged
    assert r.energy == pytest.approx(ref["energy"], abs=1e-8)
    assert r.energy == pytest.approx(pin, abs=1e-7)


def test_rhf_water_dimer_screened_store():
    import jax_qc

    mol = M.water_cluster((2, 1, 1), seed=4)
    bs = build_basis_set(mol, "even-tempered")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=1e-12)
    r = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf"))
    ref = scf.run_rhf(mol, mmd.compute_integrals(mol, bs))
    assert r.converged and ref["converged"]
    assert r.energy == pytest.approx(ref["energy"], abs=1e-8)


@pytest.mark.parametrize("mk", [M.oh_radical, M.o2_triplet], ids=["OH", "O2"])
def test_uhf_open_shell(mk):
    import jax_qc

    mol = mk()
    bs = build_basis_set(mol, "sto-3g")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=0.0)
    r = jax_qc.run_scf(mol, ints, CalcConfig(method="uhf"))
    ref = scf.run_uhf(mol, mmd.compute_integrals(mol, bs))
    assert r.converged and ref["converged"]
    assert r.energy == pytest.approx(ref["energy"], abs=1e-8)


def test_uhf_closed_shell_equals_rhf_on_gpu():
    import jax_qc

    mol = M.h2o_tier2()
    bs = build_basis_set(mol, "6-31g")
    ints = jax_qc.compute_integrals(mol, bs, eri="store", tau=0.0)
    a = jax_qc.run_scf(mol, ints, CalcConfig(method="uhf"))
    b = jax_qc.run_scf(mol, ints, CalcConfig(method="rhf"))
    assert a.energy == pytest.approx(b.energy, abs=1e-8)


def test_fock_of_zero_density_is_hcore():
    import jax_qc
    from jax_qc.scf import build_fock_rhf, build_fock_uhf

    mol = M.h2o_tier2()
    bs = build_basis_set(mol, "sto-3g")
    for eri in ("dense", "store"):
        ints = jax_qc.compute_integrals(mol, bs, eri=eri, tau=0.0)
        Z = torch.zeros_like(ints.S)
        assert torch.allclose(build_fock_rhf(ints.H_core, Z, ints.ERI), ints.H_core, atol=1e-14)
        Fa, Fb = build_fock_uhf(ints.H_core, Z, Z, ints.ERI)
        assert torch.allclose(Fa, ints.H_core, atol=1e-14)


def test_water8_energy_rotation_invariant():
    """Full-size (H2O)_8 / even-tempered: the RHF energy is invariant under a
    rigid rotation + translation of the cluster (size-independent check of
    every integral class and the supermatrix Fock build)."""
    import jax_qc
    from jax_qc.core.types import make_molecule

    mol = M.water_cluster((2, 2, 2), seed=0)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    mol2 = make_molecule(mol.elements, mol.coords @ R.T + np.array([0.3, -1.1, 2.0]),
                         mol.atomic_numbers)
    E = []
    for m in (mol, mol2):
        bs = build_basis_set(m, "even-tempered")
        ints = jax_qc.compute_integrals(m, bs, eri="store", tau=1e-12)
        r = jax_qc.run_scf(m, ints, CalcConfig(method="rhf", scf_conv=1e-10))
        assert r.converged
        E.append(r.energy)
        del ints
        torch.cuda.empty_cache()
    assert E[0] == pytest.approx(E[1], abs=1e-8)


def test_purification_graph_replays_match_eager():
    """The graph-replayed TC2 iteration gives the eager iteration's
    projectors on a sequence of Fock pairs (same shape, so the captured
    graphs are replayed for each), with eager work between the calls, and
    reports the same non-convergence for a spectrum without a gap."""
    import numpy as np

    from jax_qc.scf import purify

    rng = np.random.default_rng(11)
    n = 96
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    for rep in range(4):
        Fs = []
        for b in range(2):
            w = np.sort(rng.uniform(-20, 5, n))
            w[17 + b:] += 0.5  # a gap after 17 (alpha) / 18 (beta) occupied
            Fs.append((Q * w[None, :]) @ Q.T)
        F = torch.from_numpy(np.stack(Fs)).cuda()
        try:
            purify.USE_GRAPHS = True
            Pg = purify.occupied_projectors(F, [17, 18])
            _ = (F @ F).sum().item()  # eager work between replays
            purify.USE_GRAPHS = False
            Pe = purify.occupied_projectors(F, [17, 18])
        finally:
            purify.USE_GRAPHS = True
        assert Pg is not None and Pe is not None
        assert float((Pg - Pe).abs().max()) < 1e-13, rep
    w = np.linspace(-10.0, 10.0, n)
    w[5] = w[4]
    F = torch.from_numpy((Q * w[None, :]) @ Q.T)[None].cuda()
    assert purify.occupied_projectors(F, [5], max_iter=80) is None
