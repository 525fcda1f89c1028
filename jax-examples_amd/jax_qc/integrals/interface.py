"""compute_integrals: every matrix an SCF needs (replaces
jax_qc/integrals/interface.py:26-67), evaluated on the GPU.

``eri='dense'`` returns the reference's dense (n,n,n,n) tensor; ``'store'``
returns an :class:`ERIStore` (unique integrals as packed supermatrices in
HBM), which ``build_fock_rhf/uhf`` contract without ever forming n^4
elements; ``'auto'`` picks dense up to 48 functions.
"""

from __future__ import annotations

import time
from typing import Optional

from jax_qc import _native as nat
from jax_qc.core.types import IntegralSet
from jax_qc.integrals import engine as eng
from jax_qc.integrals.eri import compute_eri_tensor
from jax_qc.integrals.one_electron import (
    compute_kinetic_matrix,
    compute_nuclear_matrix,
    compute_overlap_matrix,
    nuclear_repulsion_energy,
)
from jax_qc.integrals.store import ERIStore
from jax_qc.profiling import optional_stage


def compute_integrals(mol, basis, timer=None, eri: str = "auto", tau: Optional[float] = None,
                      tau32: Optional[float] = None, prim_eps: Optional[float] = None) -> IntegralSet:
    """tau: Schwarz threshold; tau32: quartets with a bound below it sum
    their primitives in FP32; prim_eps: the ERI store leaves out primitive
    pairs whose Cauchy-Schwarz bound against the largest one is below it
    (engine.screen_primitives; mixed path only)."""
    dev = nat.device()
    nat.ensure_init(dev)
    if eri == "auto":
        eri = "dense" if basis.n_basis <= 48 else "store"
    with optional_stage(timer, "integrals"):
        pt = eng.PairTable(basis, dev)
        with optional_stage(timer, "overlap"):
            S = compute_overlap_matrix(basis, pt)
        with optional_stage(timer, "kinetic"):
            T = compute_kinetic_matrix(basis, pt)
        with optional_stage(timer, "nuclear"):
            V = compute_nuclear_matrix(basis, mol, pt)
        with optional_stage(timer, "h_core"):
            H = T + V
        with optional_stage(timer, "eri"):
            if eri == "dense":
                ERI = compute_eri_tensor(basis, 0.0 if tau is None else tau, tau32, pt)
            elif eri == "store":
                ept = eng.screen_primitives(pt, prim_eps) if prim_eps else pt
                ERI = ERIStore(basis, 1e-12 if tau is None else tau, tau32, dev, ept)
            else:
                raise ValueError(f"eri must be 'auto', 'dense' or 'store', got {eri!r}")
        with optional_stage(timer, "nuclear_repulsion"):
            E_nuc = nuclear_repulsion_energy(mol)
    return IntegralSet(S=S, T=T, V=V, H_core=H, ERI=ERI, E_nuc=float(E_nuc))
