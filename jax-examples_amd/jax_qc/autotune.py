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
# Measured on MI355X, (H2O)_8 even-tempered, FP64, per launch (one thread vs
# one wave per quartet): per-thread wins for (dp|pp) 14.9 vs 78 ms, (dp|ds)
# 4.3 vs 58, (dd|ps) 3.1 vs 13.8, (dd|ds) 6.3 vs 11.3; the wave kernel wins
# for (dp|dp) 14.7 vs 55, (dd|pp) 7.2 vs 68, (dd|dp) 7.5 vs 124, (dd|dd) 1.5
# vs 328.  Lower classes: per-thread (by 5-40x).
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


# --- batched SCF: work-group shape per quartet class, Fock segment length ---

# (variant, quartets per work-group): one thread per quartet in 64- or
# 128-lane work-groups, or one wave per quartet
BATCH_ERI_CHOICES = ((1, 64), (1, 128), (2, 0))
BATCH_SEGS = (1, 2, 4, 8, 16)


def autotune_batch(b, repeats: int = 2, eri_choices=BATCH_ERI_CHOICES, segs=BATCH_SEGS,
                   verbose: bool = False) -> dict:
    """Time every class launch of a ``BatchedRHF``'s quartet plan with each
    (kernel, work-group) choice and the batched Fock build with each segment
    length; returns {"eri": {(cab, ccd): (variant, block)}, "seg": s,
    "timings": ...} for ``BatchedRHF(eri_tuning=..., seg=...)``.  A class
    pair with FP64 and FP32 launches gets the choice with the smaller sum."""
    J = torch.zeros(int(b.bpt.sm_base_m[-1]), dtype=torch.float64, device=b.dev)
    per_class: Dict[Tuple[int, int], Dict[Tuple[int, int], float]] = {}
    timings = {}
    for L in b.plan.launches:
        res = {}
        for v, blk in eri_choices:
            b.eri_launch(L, J, v, blk)  # warm (code object load)
            torch.cuda.synchronize(b.dev)
            t = time.perf_counter()
            for _ in range(repeats):
                b.eri_launch(L, J, v, blk)
            torch.cuda.synchronize(b.dev)
            res[(v, blk)] = (time.perf_counter() - t) / repeats * 1e3
        timings[(L.cab, L.ccd, L.fp32)] = res
        acc = per_class.setdefault((L.cab, L.ccd), {})
        for k, ms in res.items():
            acc[k] = acc.get(k, 0.0) + ms
        if verbose:
            print(f"autotune batch {(L.cab, L.ccd, L.fp32)} nq={L.nquartets}: "
                  + ", ".join(f"v{v}/{blk}={ms:.3f} ms" for (v, blk), ms in res.items()), flush=True)
    del J
    eri = {k: min(v, key=v.get) for k, v in per_class.items()}
    seg_ms = {}
    D = torch.zeros(b.nmol, b.nmax, b.nmax, dtype=torch.float64, device=b.dev)
    for sg in segs:
        b.set_seg(sg)
        b.fock(D)  # captures
        torch.cuda.synchronize(b.dev)
        t = time.perf_counter()
        for _ in range(max(5, repeats)):
            b.graph.replay()
        torch.cuda.synchronize(b.dev)
        seg_ms[sg] = (time.perf_counter() - t) / max(5, repeats) * 1e3
    seg = min(seg_ms, key=seg_ms.get)
    b.set_seg(seg)
    if verbose:
        print("autotune batch seg: " + ", ".join(f"{k}={v:.3f} ms" for k, v in seg_ms.items()),
              flush=True)
    return {"eri": eri, "seg": seg, "timings": {"eri": timings, "seg": seg_ms}}


def save_batch(tuning: dict, path: str) -> None:
    with open(path, "w") as fh:
        json.dump({"eri": {f"{a},{c}": list(v) for (a, c), v in tuning["eri"].items()},
                   "seg": tuning["seg"]}, fh, indent=1)


def load_batch(path: str) -> dict:
    with open(path) as fh:
        d = json.load(fh)
    return {"eri": {tuple(int(x) for x in k.split(",")): tuple(v) for k, v in d["eri"].items()},
            "seg": int(d["seg"])}


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
