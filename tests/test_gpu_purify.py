"""Trace-correcting purification on the device (graph replays of batched
steps) against eigh projectors."""

import pytest
import torch

pytestmark = pytest.mark.gpu


def _spectrum_matrix(n, n_occ, gap, seed):
    """Symmetric n x n matrix with eigenvalues spread over [-1, 1] and a gap
    of `gap` between the n_occ-th and the next one; returns (F, projector
    onto the n_occ lowest eigenvectors)."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    Q, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64))
    lo = torch.linspace(-1.0, -gap / 2, n_occ, dtype=torch.float64)
    hi = torch.linspace(gap / 2, 1.0, n - n_occ, dtype=torch.float64)
    lam = torch.cat([lo, hi])
    F = (Q * lam) @ Q.T
    P = Q[:, :n_occ] @ Q[:, :n_occ].T
    return 0.5 * (F + F.T), P


@pytest.mark.parametrize("fused", [True, False], ids=["fused-step", "three-launch-step"])
@pytest.mark.parametrize("n,n_occ", [(192, 70), (200, 81), (75, 30)])
def test_converged_matrices_stay_converged_across_replayed_chunks(fused, n, n_occ):
    """A call that needs many step chunks (gap 1e-4 of the spectral width)
    followed by one that converges in a few (wide gap), same shapes: the
    second call replays the first call's chunk count before it reads the
    convergence measure, i.e. steps past convergence; qc_tc2_update and
    qc_tc2_step leave a converged matrix unchanged (without that, its
    rounding grows by 2x per step and the projector is lost), so every
    projector matches eigh.  Both step forms; n = 200 and 75 leave partial
    32 x 32 tiles (75: odd rows, no 16-byte loads)."""
    from jax_qc.scf import purify

    dev = torch.device("cuda", 0)
    purify.FUSED_STEP = fused
    try:
        for gap, seed in ((1e-4, 1), (0.5, 2), (1e-4, 3), (0.3, 4)):
            F, P = _spectrum_matrix(n, n_occ, gap, seed)
            F2 = torch.stack([F, F.flip(0).flip(1)]).to(dev)  # same spectrum, permuted basis
            P2 = torch.stack([P, P.flip(0).flip(1)])
            got = purify.occupied_projectors(F2, [n_occ, n_occ])
            assert got is not None, gap
            assert float((got.cpu() - P2).abs().max()) < 1e-9, gap
    finally:
        purify.FUSED_STEP = True
