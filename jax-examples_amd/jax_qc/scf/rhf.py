"""Restricted Hartree-Fock driver (replaces jax_qc/scf/rhf.py:52-169).

Same step order and convergence rule as the reference: Fock build ->
optional damping -> DIIS -> diagonalise in the X basis -> density ->
energy; converged when |dE| < scf_conv and max|e| < max(scf_conv, 1e-6).
All matrices are FP64 device tensors; the Fock build is the hot path.  The
density of a step is the occupied projector of X^T F X, reached by
purification (jax_qc.scf.purify; eigh when it does not converge); the
converged state's orbitals come from one eigh of its Fock matrix.
"""

from __future__ import annotations

import dataclasses
from typing import Optional

import torch

from jax_qc.core.types import CalcConfig, IntegralSet, Molecule, SCFResult, SCFState
from jax_qc.profiling import optional_stage
from jax_qc.scf.damping import damp
from jax_qc.scf.density import density_rhf
from jax_qc.scf.diis import diis_extrapolate, diis_history_init
from jax_qc.scf.energy import electronic_energy_rhf
from jax_qc.scf.fock import build_fock_rhf, last_build_was_full
from jax_qc.scf.guess import core_guess
from jax_qc.scf.orthogonalize import symmetric_orthogonalization
from jax_qc.scf import purify


def _diagonalize_fock(F, X):
    e, Cp = torch.linalg.eigh(X.T @ F @ X)
    return e, X @ Cp


def _density_step(F, X, n_occ):
    """(D, orbital energies, C) of one Roothaan step; the last two are None
    when the density came from purification."""
    if purify.ENABLED:
        Xt = X.T.contiguous()
        P = purify.occupied_projectors((Xt @ F @ X)[None], [n_occ])
        if P is not None:
            return 2.0 * (X @ P[0] @ Xt), None, None
    orb_e, C = _diagonalize_fock(F, X)
    return density_rhf(C, n_occ), orb_e, C


def scf_step_rhf(state, H_core, ERI, S, X, n_occ, diis_history, damping, timer=None):
    with optional_stage(timer, "fock_build"):
        F_new = build_fock_rhf(H_core, state.density, ERI)
    with optional_stage(timer, "damp"):
        F_d = damp(F_new, state.fock, damping) if damping > 0.0 else F_new
    with optional_stage(timer, "diis"):
        F_ext, diis_history, error = diis_extrapolate(diis_history, F_d, state.density, S, X)
    with optional_stage(timer, "diag"):
        D_new, orb_e, C = _density_step(F_ext, X, n_occ)
    with optional_stage(timer, "energy"):
        E = float(electronic_energy_rhf(D_new, H_core, F_ext))
    new = SCFState(density=D_new, fock=F_ext, coefficients=C, orbital_energies=orb_e, energy=E,
                   iteration=state.iteration + 1)
    return new, diis_history, error


def run_rhf(mol: Molecule, ints: IntegralSet, config: CalcConfig = CalcConfig(),
            timer=None) -> SCFResult:
    if mol.spin != 0:
        raise ValueError(f"run_rhf requires a closed-shell molecule (spin=0); got spin={mol.spin}.")
    if mol.n_electrons % 2 != 0:
        raise ValueError(f"run_rhf requires an even number of electrons; got {mol.n_electrons}.")
    n_occ = mol.n_electrons // 2
    H, S = ints.H_core, ints.S
    with optional_stage(timer, "scf"):
        with optional_stage(timer, "orthogonalize"):
            X = symmetric_orthogonalization(S)
        with optional_stage(timer, "initial_guess"):
            D0 = core_guess(H, X, n_occ)
            F0 = build_fock_rhf(H, D0, ints.ERI)
            E0 = float(electronic_energy_rhf(D0, H, F0))
        # no orbitals yet: the first step replaces them, and a run that takes
        # no step gets those of F0 from the final diagonalisation below
        state = SCFState(density=D0, fock=F0, coefficients=None, orbital_energies=None, energy=E0,
                         iteration=0)
        hist = diis_history_init(S.shape[0], config.diis_space)
        converged, last = False, state.energy
        for _ in range(config.max_scf_iter):
            state, hist, err = scf_step_rhf(state, H, ints.ERI, S, X, n_occ, hist,
                                            config.damping, timer)
            dE = abs(state.energy - last)
            err_max = hist.err_max  # max |e| of this step's error, read back with DIIS
            last = state.energy
            if dE < config.scf_conv and err_max < max(config.scf_conv, 1e-6):
                if not last_build_was_full(ints.ERI):
                    continue
                converged = True
                break
        if state.coefficients is None:
            with optional_stage(timer, "diag"):
                e_f, C_f = _diagonalize_fock(state.fock, X)
            state = dataclasses.replace(state, coefficients=C_f, orbital_energies=e_f)
    return SCFResult(converged=converged, state=state, energy=state.energy + float(ints.E_nuc),
                     E_elec=state.energy, E_nuc=float(ints.E_nuc), n_iterations=state.iteration,
                     S=S, H_core=H, ERI=ints.ERI)
