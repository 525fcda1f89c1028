"""Trace-correcting purification (jax_qc.scf.purify) against the projector
that eigh gives, on the CPU: gapped spectra of several widths, a batch with
different occupations, and a spectrum without a gap at n_occ (must report
non-convergence so that the SCF step falls back to eigh)."""

import numpy as np
import pytest
import torch

from jax_qc.scf import purify


def _sym_with_spectrum(w, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((len(w), len(w))))
    return torch.from_numpy((Q * np.asarray(w)[None, :]) @ Q.T)


def _eigh_projector(F, k):
    _, V = torch.linalg.eigh(F)
    return V[:, :k] @ V[:, :k].T


@pytest.mark.parametrize("n,k,gap,spread", [(40, 7, 0.5, 30.0), (120, 37, 0.05, 50.0),
                                             (200, 100, 0.3, 5.0), (60, 1, 0.2, 10.0)])
def test_projector_matches_eigh(n, k, gap, spread):
    rng = np.random.default_rng(n + k)
    occ = -np.sort(rng.uniform(gap / 2, spread / 2, k))
    vir = np.sort(rng.uniform(gap / 2, spread / 2, n - k))
    F = _sym_with_spectrum(np.concatenate([occ, vir]), seed=k)
    P = purify.occupied_projectors(F[None], [k])
    assert P is not None
    ref = _eigh_projector(F, k)
    assert float((P[0] - ref).abs().max()) < 1e-10
    assert abs(float(torch.trace(P[0])) - k) < 1e-9
    assert float((P[0] @ P[0] - P[0]).abs().max()) < 1e-10


def test_batch_with_different_occupations():
    Fa = _sym_with_spectrum(np.linspace(-20.0, 5.0, 50), seed=1)
    Fb = _sym_with_spectrum(np.linspace(-15.0, 8.0, 50), seed=2)
    P = purify.occupied_projectors(torch.stack([Fa, Fb]), [9, 7])
    assert P is not None
    assert float((P[0] - _eigh_projector(Fa, 9)).abs().max()) < 1e-10
    assert float((P[1] - _eigh_projector(Fb, 7)).abs().max()) < 1e-10


def test_no_gap_is_reported():
    w = np.linspace(-10.0, 10.0, 30)
    w[5] = w[4]  # eigenvalues 5 and 6 (1-based) coincide: no gap after 5 occupied
    F = _sym_with_spectrum(w, seed=3)
    assert purify.occupied_projectors(F[None], [5], max_iter=80) is None
