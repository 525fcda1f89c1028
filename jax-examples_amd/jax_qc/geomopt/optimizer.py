"""BFGS geometry optimisation (restates jax_qc/geomopt/optimizer.py:79-209):
identity initial inverse Hessian, step clipped to ``max_step_size`` Bohr,
backtracking Armijo line search (c1 = 1e-4, at most 10 halvings, the last
trial accepted), convergence on max |g| < grad_tol or max |s| < step_tol
with max |g| < 10 grad_tol, update skipped when s.y <= 1e-10.

Energies and gradients come from the GPU path; unlike the reference (which
always applies its RHF gradient, optimizer.py:75) UHF runs use the UHF
gradient.  ``callback(step, coords, energy, grad, info)`` sees every accepted
point (used to compare precision modes step by step).
"""

from __future__ import annotations

import dataclasses
from typing import Callable, NamedTuple, Optional

import numpy as np
import torch

from jax_qc.basis import build_basis_set
from jax_qc.core.types import CalcConfig
from jax_qc.grad.analytic import scf_gradient
from jax_qc.grad.numerical_grad import displaced
from jax_qc.integrals import compute_integrals
from jax_qc.scf import run_scf


class OptResult(NamedTuple):
    converged: bool
    molecule: object
    energy: float
    gradient: np.ndarray
    n_steps: int
    trajectory: list


def energy_and_gradient(coords, mol, basis_name, config, eri="store", tau=1e-12, tau32=None,
                        tile_f32=None, tile_i16=0.0, tile_i8=0.0, grad_tau=1e-13,
                        grad_tau_density=0.0, grad_tau32=None,
                        prim_eps=None, fock_tau=None, fock_full_every=0, timings=None,
                        keep_store=False, grad_prim_eps=None):
    """Energy and gradient at ``coords``.  Mixed-path options: tau32 (FP32
    ERI sums below that Schwarz bound), tile_f32 (FP32 supermatrix tiles;
    with it, tile_i16 / tile_i8: int16 / int8 tiles with a per-tile scale),
    prim_eps (primitive-pair screening of the ERIs and derivative ERIs;
    grad_prim_eps, when set, screens the derivative ERIs' primitive pairs at
    its own level),
    fock_tau (incremental Fock builds with tile screening), grad_tau_density
    and grad_tau32 (gradient screening / FP32 sums).  ``timings`` (a dict)
    receives the wall time of each stage, the device synchronised between
    them.  keep_store: the returned result keeps its integral store
    (``res.ERI``) for further Fock builds."""
    import time

    def mark(name, t0):
        if timings is not None:
            torch.cuda.synchronize()
            timings[name] = time.perf_counter() - t0
        return time.perf_counter()

    t = time.perf_counter()
    m = displaced(mol, coords)
    bs = build_basis_set(m, basis_name)
    ints = compute_integrals(m, bs, eri=eri, tau=tau, tau32=tau32, prim_eps=prim_eps)
    if eri == "store":
        if config.method.lower() == "uhf":
            ints.ERI.assemble_p_k()  # UHF Fock builds from P and K (see ERIStore.fock_uhf)
        else:
            ints.ERI.ensure_P()
            ints.ERI.release_J()
        if tile_f32 is not None:
            ints.ERI.use_compact(tau_f32=tile_f32, tau_i16=tile_i16, tau_i8=tile_i8)
        if fock_tau is not None:
            ints.ERI.set_incremental(fock_tau, fock_full_every)
    t = mark("integrals", t)
    res = run_scf(m, ints, config)
    t = mark("scf", t)
    gpt = getattr(ints.ERI, "pt", None)
    if grad_prim_eps is not None and gpt is not None:
        from jax_qc.integrals.engine import screen_primitives

        gpt = screen_primitives(getattr(gpt, "unscreened", gpt), grad_prim_eps)
    g = scf_gradient(m, bs, res, grad_tau, pair_table=gpt, tau_density=grad_tau_density,
                     tau32=grad_tau32)
    t = mark("gradient", t)
    if timings is not None and hasattr(ints.ERI, "inc") and ints.ERI.inc is not None:
        timings["fock_full_builds"] = ints.ERI.inc.n_full
        timings["fock_incremental_builds"] = ints.ERI.inc.n_inc
    if keep_store:
        return float(res.energy), g.cpu().numpy(), m, res
    # drop the integral store (tens of GB at 480 functions) before the
    # caller's next evaluation
    res = dataclasses.replace(res, ERI=None)
    out = float(res.energy), g.cpu().numpy(), m, res
    del ints
    torch.cuda.empty_cache()
    return out


def optimize_bfgs(mol, basis_name="sto-3g", method="rhf", max_steps=50, grad_tol=1e-4,
                  step_tol=1e-6, scf_conv=1e-10, max_step_size=0.3, verbose=1,
                  callback: Optional[Callable] = None, **eg_opts) -> OptResult:
    config = CalcConfig(method=method, basis=basis_name, scf_conv=scf_conv, max_scf_iter=200)
    coords = np.asarray(mol.coords, dtype=np.float64)
    n = coords.size
    energy, grad, cur, res = energy_and_gradient(coords, mol, basis_name, config, **eg_opts)
    g = grad.reshape(-1)
    Hinv = np.eye(n)
    traj = [(coords.copy(), energy)]
    converged = False
    if callback:
        callback(0, coords, energy, grad, res)
    if verbose:
        print(f"  Step   0: E = {energy:16.10f}  |g|_max = {np.abs(g).max():.2e}", flush=True)
    for step in range(1, max_steps + 1):
        gmax = float(np.abs(g).max())
        if gmax < grad_tol:
            converged = True
            break
        d = -Hinv @ g
        nrm = float(np.linalg.norm(d))
        if nrm > max_step_size:
            d *= max_step_size / nrm
        alpha, e_old, g_old = 1.0, energy, g.copy()
        for _ in range(10):
            new = coords + alpha * d.reshape(coords.shape)
            e_new, g_new, m_new, res = energy_and_gradient(new, mol, basis_name, config, **eg_opts)
            if e_new <= e_old + 1e-4 * alpha * float(g_old @ d):
                break
            alpha *= 0.5
        s = alpha * d
        coords, energy, cur = new, e_new, m_new
        gn = g_new.reshape(-1)
        y = gn - g_old
        smax = float(np.abs(s).max())
        traj.append((coords.copy(), energy))
        if callback:
            callback(step, coords, energy, g_new, res)
        if smax < step_tol and gmax < grad_tol * 10:
            g = gn
            converged = True
            break
        sy = float(s @ y)
        if sy > 1e-10:
            rho = 1.0 / sy
            V = np.eye(n) - rho * np.outer(s, y)
            Hinv = V @ Hinv @ V.T + rho * np.outer(s, s)
        g = gn
        if verbose:
            print(f"  Step {step:3d}: E = {energy:16.10f}  |g|_max = {np.abs(g).max():.2e}  "
                  f"|s|_max = {smax:.2e}", flush=True)
    return OptResult(converged, cur, energy, g.reshape(coords.shape), len(traj) - 1, traj)
