"""Unrestricted Hartree-Fock driver (replaces jax_qc/scf/uhf.py:187-287).

Combined (2n, n) DIIS, damping 0.3 for the first 5 iterations of open-shell
runs, the alpha-repulsion + perturbation guess, and the reference's
convergence rule.  The result packs the total density and alpha orbitals
into ``state`` as the reference does, and additionally keeps both spin
channels in ``SCFResult.uhf`` for the UHF gradient.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass

import torch

from jax_qc.core.types import CalcConfig, IntegralSet, Molecule, SCFResult, SCFState
from jax_qc.profiling import optional_stage
from jax_qc.scf.damping import damp
from jax_qc.scf.density import density_uhf
from jax_qc.scf.diis import diis_extrapolate_uhf, diis_history_init_uhf
from jax_qc.scf.energy import electronic_energy_uhf
from jax_qc.scf.fock import build_fock_uhf, last_build_was_full
from jax_qc.scf.guess import core_guess_uhf
from jax_qc.scf.orthogonalize import symmetric_orthogonalization
from jax_qc.scf import purify


def _diagonalize_pair(Fa, Fb, X):
    """Both spins' Roothaan problems in one batched eigensolver call (the
    solver's cost at a few hundred functions is latency, not arithmetic, so
    two matrices cost about what one does)."""
    Xt = X.T.contiguous()
    e, Cp = torch.linalg.eigh(Xt @ torch.stack([Fa, Fb]) @ X)
    return (e[0], X @ Cp[0]), (e[1], X @ Cp[1])


def _density_pair_step(Fa, Fb, X, n_alpha, n_beta):
    """(Da, Db, orbitals) of one UHF step; orbitals is ((e_a, C_a), (e_b,
    C_b)), or None when the densities came from purification."""
    if purify.ENABLED:
        Xt = X.T.contiguous()
        P = purify.occupied_projectors(Xt @ torch.stack([Fa, Fb]) @ X, [n_alpha, n_beta])
        if P is not None:
            D = X @ P @ Xt
            return D[0], D[1], None
    (oa, Ca), (ob, Cb) = _diagonalize_pair(Fa, Fb, X)
    Da, Db = density_uhf(Ca, Cb, n_alpha, n_beta)
    return Da, Db, ((oa, Ca), (ob, Cb))


@dataclass
class UHFState:
    density_alpha: torch.Tensor
    density_beta: torch.Tensor
    fock_alpha: torch.Tensor
    fock_beta: torch.Tensor
    coefficients_alpha: torch.Tensor
    coefficients_beta: torch.Tensor
    orbital_energies_alpha: torch.Tensor
    orbital_energies_beta: torch.Tensor
    energy: float
    iteration: int


def scf_step_uhf(state, H, ERI, S, X, n_alpha, n_beta, hist, damping, timer=None):
    with optional_stage(timer, "fock_build"):
        Fa, Fb = build_fock_uhf(H, state.density_alpha, state.density_beta, ERI)
    with optional_stage(timer, "damp"):
        if damping > 0.0:
            Fa = damp(Fa, state.fock_alpha, damping)
            Fb = damp(Fb, state.fock_beta, damping)
    with optional_stage(timer, "diis"):
        Fa, Fb, hist, ea, eb = diis_extrapolate_uhf(hist, Fa, Fb, state.density_alpha,
                                                    state.density_beta, S, X)
    with optional_stage(timer, "diag"):
        Da, Db, orb = _density_pair_step(Fa, Fb, X, n_alpha, n_beta)
    (oa, Ca), (ob, Cb) = orb if orb is not None else ((None, None), (None, None))
    with optional_stage(timer, "energy"):
        E = float(electronic_energy_uhf(Da, Db, H, Fa, Fb))
    err = hist.err_max  # max |e| over both spins (the stacked error of the last push)
    new = UHFState(Da, Db, Fa, Fb, Ca, Cb, oa, ob, E, state.iteration + 1)
    return new, hist, err


def run_uhf(mol: Molecule, ints: IntegralSet, config: CalcConfig = CalcConfig(),
            timer=None) -> SCFResult:
    na, nb = mol.n_alpha, mol.n_beta
    H, S = ints.H_core, ints.S
    with optional_stage(timer, "scf"):
        with optional_stage(timer, "orthogonalize"):
            X = symmetric_orthogonalization(S)
        with optional_stage(timer, "initial_guess"):
            Da, Db, _ = core_guess_uhf(H, X, na, nb, ints.ERI, orbitals=False)
            Fa, Fb = build_fock_uhf(H, Da, Db, ints.ERI)
            Da, Db, orb = _density_pair_step(Fa, Fb, X, na, nb)
            (oa, Ca), (ob, Cb) = orb if orb is not None else ((None, None), (None, None))
            E0 = float(electronic_energy_uhf(Da, Db, H, Fa, Fb))
        state = UHFState(Da, Db, Fa, Fb, Ca, Cb, oa, ob, E0, 0)
        hist = diis_history_init_uhf(S.shape[0], config.diis_space)
        init_damp = 0.3 if na != nb else 0.0
        converged, last = False, state.energy
        for it in range(config.max_scf_iter):
            dv = init_damp if it < 5 else config.damping
            state, hist, err = scf_step_uhf(state, H, ints.ERI, S, X, na, nb, hist, dv, timer)
            dE = abs(state.energy - last)
            last = state.energy
            if dE < config.scf_conv and err < max(config.scf_conv, 1e-6):
                if not last_build_was_full(ints.ERI):
                    continue
                converged = True
                break
        if state.coefficients_alpha is None:
            with optional_stage(timer, "diag"):
                (oa, Ca), (ob, Cb) = _diagonalize_pair(state.fock_alpha, state.fock_beta, X)
            state = dataclasses.replace(state, coefficients_alpha=Ca, coefficients_beta=Cb,
                                        orbital_energies_alpha=oa, orbital_energies_beta=ob)
    total = state.density_alpha + state.density_beta
    st = SCFState(density=total, fock=state.fock_alpha, coefficients=state.coefficients_alpha,
                  orbital_energies=state.orbital_energies_alpha, energy=state.energy,
                  iteration=state.iteration)
    return SCFResult(converged=converged, state=st, energy=state.energy + float(ints.E_nuc),
                     E_elec=state.energy, E_nuc=float(ints.E_nuc), n_iterations=state.iteration,
                     S=S, H_core=H, ERI=ints.ERI, uhf=state)
