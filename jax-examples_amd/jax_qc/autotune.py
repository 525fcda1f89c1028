"""Per-class kernel selection for the ERI class kernels.

Each (bra pair class, ket pair class, precision) has two kernels: one thread
per contracted quartet (intermediates in VGPRs; fast while they fit) and one
64-lane wave per quartet (R_tuv and E tables in LDS; no spills for the high
classes).  ``autotune_eri`` times both on the real launches of a quartet
plan and records the faster in ``engine.ERI_VARIANT``; ``DEFAULT_VARIANTS``
holds the choice measured on MI355X for the (H2O)_8 even-tempered workload
(kernel trace of bench.py), used until a tuning run replaces it.
"""

from __future__ import annotations

import json
import os
import time
from typing import Dict, Tuple

import torch

from jax_qc.integrals import engine as eng

# (cab, ccd) -> variant, pair classes ss=0 ps=1 pp=2 ds=3 dp=4 dd=5.
# Measured (FP64, per launch): per-thread wins for (dp|pp) 15.7 vs 141 ms,
# (dp|ds) 4.5 vs 62, (dd|ps) 3.2 vs 29, (dd|ds) 6.2 vs 13; the wave kernel
# wins for (dp|dp) 24 vs 46, (dd|pp) 31 vs 69, (dd|dp) 35 vs 124,
# (dd|dd) 13 vs 322.  Lower classes: per-thread.
DEFAULT_VARIANTS: Dict[Tuple[int, int], int] = {
    (4, 2): 1, (4, 3): 1, (5, 1): 1, (5, 3): 1,
    (4, 4): 2, (5, 2): 2, (5, 4): 2, (5, 5): 2,
}


def use_defaults() -> None:
    eng.ERI_VARIANT.update(DEFAULT_VARIANTS)


def autotune_eri(pt, plan, out: torch.Tensor | None = None, mode: int = eng.MODE_SUPER,
                 repeats: int = 2, variants=(1, 2), min_quartets: int = 1,
                 verbose: bool = False) -> Dict[Tuple[int, int, int], Dict[int, float]]:
    """Time every launch of ``plan`` with each variant; keep the fastest
    per (cab, ccd) in engine.ERI_VARIANT.  Returns the timings (ms)."""
    if out is None:
        out = torch.zeros(eng.sm_numel(pt.nbf), dtype=torch.float64, device=pt.dev)
    timings = {}
    saved = eng.FORCE_VARIANT
    try:
        for L in plan.launches:
            if L.nquartets < min_quartets:
                continue
            key = (L.cab, L.ccd, L.fp32)
            sub = eng.QuartetPlan(launches=[L])
            res = {}
            for v in variants:
                eng.FORCE_VARIANT = v
                eng.run_eri(pt, sub, mode, out)  # warm (code object load)
                torch.cuda.synchronize(pt.dev)
                t = time.perf_counter()
                for _ in range(repeats):
                    eng.run_eri(pt, sub, mode, out)
                torch.cuda.synchronize(pt.dev)
                res[v] = (time.perf_counter() - t) / repeats * 1e3
            timings[key] = res
            best = min(res, key=res.get)
            eng.ERI_VARIANT[(L.cab, L.ccd)] = best
            if verbose:
                print(f"autotune {key}: " + ", ".join(f"v{v}={ms:.3f} ms" for v, ms in res.items()),
                      flush=True)
    finally:
        eng.FORCE_VARIANT = saved
    return timings


def save(path: str) -> None:
    with open(path, "w") as fh:
        json.dump({f"{a},{b}": v for (a, b), v in eng.ERI_VARIANT.items()}, fh, indent=1)


def load(path: str) -> None:
    if os.path.exists(path):
        with open(path) as fh:
            for k, v in json.load(fh).items():
                a, b = (int(x) for x in k.split(","))
                eng.ERI_VARIANT[(a, b)] = int(v)


use_defaults()
