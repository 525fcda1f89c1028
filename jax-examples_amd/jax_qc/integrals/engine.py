"""Host planning for the device integral kernels.

* ``PairTable``: shell pairs (i >= j in basis order, each oriented so the
  first shell has the larger l) grouped by pair class la(la+1)/2 + lb, and the
  primitive-pair records (p, P, P-A, P-B, K, alpha) the kernels read.  This
  is the per-pair data the reference recomputes inside every primitive
  quartet (jax_qc/integrals/obara_saika.py:237-250).
* ``schwarz``: Q_pair = sqrt(max_ab (ab|ab)) on the device (screening.py:32-58).
* ``QuartetPlan``: for each class pair (bra class >= ket class) the surviving
  quartets after Schwarz screening Q_bra Q_ket >= tau, without materialising
  quartet lists: pairs are sorted by Q (descending) within each class, so the
  survivors of a bra are a prefix of the ket list; the kernel maps a global
  index to (bra, ket) through the per-bra prefix sums.  With a precision split
  the FP64 quartets (bound >= tau32) are the first k64 kets of each bra and
  the FP32 quartets the rest.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np
import torch

from jax_qc import _native as nat
from jax_qc.basis.build import shell_arrays

MODE_DENSE, MODE_SUPER, MODE_SCHWARZ = 0, 1, 2
SM_B = 64

# Kernel variant per (bra class, ket class): 0 = the library's built-in
# choice, 1 = one thread per quartet, 2 = one wave per quartet.  Filled in
# by jax_qc.autotune (or by hand); ``FORCE_VARIANT`` overrides every class.
ERI_VARIANT: Dict[Tuple[int, int], int] = {}
FORCE_VARIANT: int | None = None


def variant_for(cab: int, ccd: int) -> int:
    if FORCE_VARIANT is not None:
        return FORCE_VARIANT
    return ERI_VARIANT.get((cab, ccd), 0)


def pair_class(la: int, lb: int) -> int:
    return la * (la + 1) // 2 + lb


CLASS_L = {0: (0, 0), 1: (1, 0), 2: (1, 1), 3: (2, 0), 4: (2, 1), 5: (2, 2)}


def shell_pair_records(sa, a: int, b: int, prim_cut: float = 0.0) -> np.ndarray:
    """Primitive-pair records of shells (a, b) (l_a >= l_b) from
    ``shell_arrays`` output: [p, P(3), P-A(3), P-B(3), K, alpha] per row, K
    holding both contraction coefficients and exp(-alpha beta/p |A-B|^2)
    (gaussian_product.py:23-45, obara_saika.py:56-65)."""
    ex, co, ce, off, npr = (sa[k] for k in ("exps", "coefs", "centers", "prim_off", "nprim"))
    A, B = ce[a], ce[b]
    ea = ex[off[a]:off[a] + npr[a]]
    eb = ex[off[b]:off[b] + npr[b]]
    ca = co[off[a]:off[a] + npr[a]]
    cb = co[off[b]:off[b] + npr[b]]
    al = np.repeat(ea, len(eb))
    be = np.tile(eb, len(ea))
    p = al + be
    P = (al[:, None] * A + be[:, None] * B) / p[:, None]
    AB2 = float(np.dot(A - B, A - B))
    K = np.repeat(ca, len(cb)) * np.tile(cb, len(ca)) * np.exp(-al * be / p * AB2)
    rec = np.empty((len(p), 12))
    rec[:, 0] = p
    rec[:, 1:4] = P
    rec[:, 4:7] = P - A
    rec[:, 7:10] = P - B
    rec[:, 10] = K
    rec[:, 11] = al
    if prim_cut > 0.0:
        rec = rec[np.abs(K) >= prim_cut]
    return rec


def all_pair_records(sa, pa, pb, prim_cut: float = 0.0):
    """``shell_pair_records`` for many shell pairs (pa[k], pb[k]) at once:
    returns the concatenated records (pair by pair, primitives of pa outer)
    and the per-pair record counts."""
    ex, co, ce, off, npr = (sa[k] for k in ("exps", "coefs", "centers", "prim_off", "nprim"))
    na, nb = npr[pa].astype(np.int64), npr[pb].astype(np.int64)
    cnt = na * nb
    pid = np.repeat(np.arange(len(pa)), cnt)
    start = np.zeros(len(pa), dtype=np.int64)
    start[1:] = np.cumsum(cnt)[:-1]
    local = np.arange(int(cnt.sum()), dtype=np.int64) - start[pid]
    ka = off[pa][pid] + local // nb[pid]
    kb = off[pb][pid] + local % nb[pid]
    al, be = ex[ka], ex[kb]
    A, B = ce[pa][pid], ce[pb][pid]
    p = al + be
    P = (al[:, None] * A + be[:, None] * B) / p[:, None]
    AB2 = np.sum((A - B) ** 2, axis=1)
    K = co[ka] * co[kb] * np.exp(-al * be / p * AB2)
    rec = np.empty((len(p), 12))
    rec[:, 0] = p
    rec[:, 1:4] = P
    rec[:, 4:7] = P - A
    rec[:, 7:10] = P - B
    rec[:, 10] = K
    rec[:, 11] = al
    if prim_cut > 0.0:
        keep = np.abs(K) >= prim_cut
        rec = rec[keep]
        cnt = np.bincount(pid[keep], minlength=len(pa)).astype(np.int64)
    return rec, cnt


def pair_records_device(sa, pa, pb, dev, prim_cut: float = 0.0):
    """``all_pair_records`` computed with torch on ``dev`` (the same
    arithmetic, element for element; on the GPU the ~n_shell^2/2 x 9 rows cost
    about a millisecond instead of ~0.1 s of host NumPy at 256 shells).
    Returns the records as a (n, 12) FP64 tensor on ``dev`` and the per-pair
    record counts (host)."""
    f64 = dict(dtype=torch.float64, device=dev)
    i64 = dict(dtype=torch.int64, device=dev)
    ex = torch.as_tensor(np.asarray(sa["exps"], dtype=np.float64), **f64)
    co = torch.as_tensor(np.asarray(sa["coefs"], dtype=np.float64), **f64)
    ce = torch.as_tensor(np.asarray(sa["centers"], dtype=np.float64), **f64)
    npr_h = np.asarray(sa["nprim"], dtype=np.int64)
    off = torch.as_tensor(np.asarray(sa["prim_off"], dtype=np.int64), **i64)
    npr = torch.as_tensor(npr_h, **i64)
    cnt_h = npr_h[pa] * npr_h[pb]
    total = int(cnt_h.sum())
    pa_t, pb_t = torch.as_tensor(pa, **i64), torch.as_tensor(pb, **i64)
    nb = npr[pb_t]
    cnt = torch.as_tensor(cnt_h, **i64)
    pid = torch.repeat_interleave(torch.arange(len(pa), **i64), cnt, output_size=total)
    start = torch.cumsum(cnt, 0) - cnt
    local = torch.arange(total, **i64) - start[pid]
    ka = off[pa_t][pid] + torch.div(local, nb[pid], rounding_mode="floor")
    kb = off[pb_t][pid] + torch.remainder(local, nb[pid])
    al, be = ex[ka], ex[kb]
    A, B = ce[pa_t][pid], ce[pb_t][pid]
    p = al + be
    P = (al[:, None] * A + be[:, None] * B) / p[:, None]
    AB2 = torch.sum((A - B) ** 2, dim=1)
    K = co[ka] * co[kb] * torch.exp(-al * be / p * AB2)
    rec = torch.empty((total, 12), **f64)
    rec[:, 0] = p
    rec[:, 1:4] = P
    rec[:, 4:7] = P - A
    rec[:, 7:10] = P - B
    rec[:, 10] = K
    rec[:, 11] = al
    if prim_cut > 0.0:
        keep = K.abs() >= prim_cut
        rec = rec[keep]
        cnt_h = torch.bincount(pid[keep], minlength=len(pa)).cpu().numpy().astype(np.int64)
    return rec.contiguous(), cnt_h


class PairTable:
    """Shell-pair and primitive-pair tables for one basis on one device."""

    def __init__(self, basis, dev=None, fp32: bool = True, prim_cut: float = 0.0):
        if not basis.spherical:
            raise NotImplementedError("the device path implements the spherical layout")
        sa = shell_arrays(basis)
        if int(sa["l"].max(initial=0)) > 2:
            raise NotImplementedError("l > 2 shells are not supported (spherical.py:64-67)")
        self.basis = basis
        self.nbf = int(basis.n_basis)
        self.nshell = len(basis.shells)
        L, npr, off, ex, co, ce, ao = (sa[k] for k in ("l", "nprim", "prim_off", "exps", "coefs",
                                                        "centers", "ao_off"))
        ncl = (L + 1) * (L + 2) // 2
        aoc = np.zeros(self.nshell + 1, dtype=np.int64)
        aoc[1:] = np.cumsum(ncl)
        self.ncart = int(aoc[-1])
        self.cart_off = aoc
        atom = sa["atom"]
        # shell pairs i >= j in (i, j) order, oriented larger l first
        ii, jj = np.tril_indices(self.nshell)
        swap = L[ii] < L[jj]
        pa = np.where(swap, jj, ii).astype(np.int64)
        pb = np.where(swap, ii, jj).astype(np.int64)
        self.dev = dev or nat.device()
        rec, sp_n = pair_records_device(sa, pa, pb, self.dev, prim_cut)
        sp_off = np.zeros(len(pa), dtype=np.int64)
        sp_off[1:] = np.cumsum(sp_n)[:-1]
        npp = int(sp_n.sum())
        self.npairs = len(pa)
        self.sp_class = (L[pa] * (L[pa] + 1) // 2 + L[pb]).astype(np.int32)
        self.sp_shells = np.stack([pa, pb], 1).astype(np.int32)
        self.sp_ao_host = np.stack([ao[pa], ao[pb]], 1).astype(np.int32)
        self.nprim_pairs = npp
        d = self.dev
        self.pairs64 = rec
        self.pairs32 = self.pairs64.float() if fp32 else None
        self.sp_prim_off = torch.from_numpy(sp_off.astype(np.int32)).to(d)
        self.sp_nprim = torch.from_numpy(np.asarray(sp_n, dtype=np.int32)).to(d)
        self.sp_ao = torch.from_numpy(self.sp_ao_host.copy()).to(d)
        a_, b_ = self.sp_shells[:, 0], self.sp_shells[:, 1]
        self.sp_aoc = torch.from_numpy(
            np.stack([aoc[a_], aoc[b_]], 1).astype(np.int32).reshape(-1)).to(d)
        self.sp_atom = torch.from_numpy(
            np.stack([atom[a_], atom[b_]], 1).astype(np.int32).reshape(-1)).to(d)
        self.sp_diag = torch.from_numpy((a_ == b_).astype(np.int32)).to(d)
        self.class_pairs: Dict[int, np.ndarray] = {
            c: np.nonzero(self.sp_class == c)[0].astype(np.int32) for c in range(6)}
        self.Q = None  # per shell pair Schwarz bound (host numpy)
        self._keep = []

    # -- launch helpers ---------------------------------------------------
    def base_args(self, mode: int, out=None, schwarz=None) -> nat.QcEriArgs:
        a = nat.QcEriArgs()
        a.pairs64 = self.pairs64.data_ptr()
        a.pairs32 = self.pairs32.data_ptr() if self.pairs32 is not None else 0
        a.sp_prim_off = self.sp_prim_off.data_ptr()
        a.sp_nprim = self.sp_nprim.data_ptr()
        a.sp_ao = self.sp_ao.data_ptr()
        a.mode = mode
        a.nbf = self.nbf
        a.out = out.data_ptr() if out is not None else 0
        a.schwarz = schwarz.data_ptr() if schwarz is not None else 0
        return a


class _PrimitivePairView:
    """Every primitive pair of a PairTable as a one-primitive 'shell pair'
    (same records, same class), so that ``schwarz`` returns the per-primitive
    -pair bound sqrt(max_ab (ab|ab)) over its Cartesian components."""

    def __init__(self, pt: PairTable):
        n = pt.sp_nprim.cpu().numpy().astype(np.int64)
        owner = np.repeat(np.arange(pt.npairs), n)
        d = pt.dev
        self.dev, self.nbf, self.npairs = d, pt.nbf, int(n.sum())
        self.pairs64, self.pairs32 = pt.pairs64, pt.pairs32
        self.sp_prim_off = torch.arange(self.npairs, dtype=torch.int32, device=d)
        self.sp_nprim = torch.ones(self.npairs, dtype=torch.int32, device=d)
        self.sp_ao = pt.sp_ao[torch.from_numpy(owner).to(d)].contiguous()
        cls = pt.sp_class[owner]
        self.class_pairs = {c: np.nonzero(cls == c)[0].astype(np.int32) for c in range(6)}
        self.owner = owner
        self.Q = None
    base_args = PairTable.base_args


def screen_primitives(pt: PairTable, eps: float) -> PairTable:
    """A copy of ``pt`` without the primitive pairs k whose bound
    Q_k = sqrt(max (ab|ab)_k) satisfies Q_k * Q_max < eps, Q_max the largest
    primitive-pair bound of the table: by the Cauchy-Schwarz inequality of
    the Coulomb inner product every primitive quartet left out is below eps
    in magnitude (times the few-percent spread of the spherical transform).
    Every shell pair keeps at least its largest primitive pair; the
    shell-pair bounds (and so the quartet plans) are those of ``pt``.  This
    is primitive-level screening for the mixed path: the ERIs and derivative
    integrals of distant pairs of tight primitives, whose Gaussian-product
    factor exp(-ab/p |AB|^2) vanishes, are never evaluated."""
    import copy

    if pt.Q is None:
        schwarz(pt)
    view = _PrimitivePairView(pt)
    qp = schwarz(view)
    keep = qp * float(qp.max()) >= eps
    owner = view.owner
    # every shell pair keeps its largest primitive pair
    best = np.full(pt.npairs, -1.0)
    np.maximum.at(best, owner, qp)
    keep |= qp == best[owner]
    rec = pt.pairs64[torch.from_numpy(np.nonzero(keep)[0]).to(pt.dev)].contiguous()
    cnt = np.bincount(owner[keep], minlength=pt.npairs).astype(np.int64)
    off = np.zeros(pt.npairs, dtype=np.int64)
    off[1:] = np.cumsum(cnt)[:-1]
    out = copy.copy(pt)
    out.pairs64 = rec
    out.pairs32 = rec.float() if pt.pairs32 is not None else None
    out.sp_prim_off = torch.from_numpy(off.astype(np.int32)).to(pt.dev)
    out.sp_nprim = torch.from_numpy(cnt.astype(np.int32)).to(pt.dev)
    out.nprim_pairs = int(cnt.sum())
    out.prim_eps = float(eps)
    out.prim_kept = float(keep.mean())
    out.unscreened = pt  # one-electron integrals and their derivatives use every primitive
    return out


def schwarz(pt: PairTable) -> np.ndarray:
    """Per-shell-pair Q = sqrt(max_ab (ab|ab)) (Cartesian diagonal)."""
    nat.ensure_init(pt.dev)
    Qd = torch.zeros(pt.npairs, dtype=torch.float64, device=pt.dev)
    for c in range(6):
        lst = pt.class_pairs[c]
        if len(lst) == 0:
            continue
        bl = torch.from_numpy(lst).to(pt.dev)
        offs = torch.arange(len(lst) + 1, dtype=torch.int64, device=pt.dev)
        ks = torch.arange(len(lst), dtype=torch.int32, device=pt.dev)
        a = pt.base_args(MODE_SCHWARZ, schwarz=Qd)
        a.bra_list = a.ket_list = bl.data_ptr()
        a.offsets = offs.data_ptr()
        a.kstart = ks.data_ptr()
        a.nbra = len(lst)
        a.nquartets = len(lst)
        a.variant = variant_for(c, c)
        nat.call("qc_eri", c, c, 0, ctypes.byref(a), nat.stream_ptr())
    pt.Q = Qd.cpu().numpy()
    return pt.Q


@dataclass
class ClassLaunch:
    cab: int
    ccd: int
    fp32: int
    nbra: int
    nquartets: int
    bra: torch.Tensor
    ket: torch.Tensor
    offsets: torch.Tensor
    kstart: torch.Tensor


@dataclass
class QuartetPlan:
    launches: List[ClassLaunch] = field(default_factory=list)
    n64: int = 0
    n32: int = 0
    n_screened: int = 0
    tau: float = 0.0
    tau32: float = 0.0
    # surviving quartets per class pair and precision, for FLOP accounting
    counts: Dict[Tuple[int, int, int], int] = field(default_factory=dict)


def plan_quartets(pt: PairTable, tau: float = 1e-12, tau32: float | None = None) -> QuartetPlan:
    """Screened quartet launches for every class pair.  tau32=None: all
    surviving quartets in FP64; otherwise quartets with bound < tau32 go to
    the FP32 kernels."""
    if pt.Q is None:
        schwarz(pt)
    Q = pt.Q
    plan = QuartetPlan(tau=tau, tau32=tau32 or 0.0)
    sorted_lists, sorted_q = {}, {}
    for c in range(6):
        lst = pt.class_pairs[c]
        order = np.argsort(-Q[lst], kind="stable")
        sorted_lists[c] = lst[order]
        sorted_q[c] = Q[lst][order]
    total = 0
    for cab in range(6):
        for ccd in range(cab + 1):
            bl, kl = sorted_lists[cab], sorted_lists[ccd]
            if len(bl) == 0 or len(kl) == 0:
                continue
            qb, qk = sorted_q[cab], sorted_q[ccd]
            neg = -qk  # ascending
            with np.errstate(divide="ignore"):
                kmax = np.searchsorted(neg, -(tau / np.maximum(qb, 1e-300)), side="right")
                k64 = (np.searchsorted(neg, -(tau32 / np.maximum(qb, 1e-300)), side="right")
                       if tau32 else kmax)
            total += len(bl) * len(kl) if cab != ccd else len(bl) * (len(bl) + 1) // 2
            if cab == ccd:
                cap = np.arange(1, len(bl) + 1)
                kmax = np.minimum(kmax, cap)
                k64 = np.minimum(k64, cap)
            k64 = np.minimum(k64, kmax)
            for fp32, (start, stop) in ((0, (np.zeros_like(k64), k64)), (1, (k64, kmax))):
                cnt = (stop - start).astype(np.int64)
                n = int(cnt.sum())
                if n == 0:
                    continue
                offs = np.zeros(len(bl) + 1, dtype=np.int64)
                offs[1:] = np.cumsum(cnt)
                # host arrays for now; moved to the device together below
                plan.launches.append(ClassLaunch(
                    cab, ccd, fp32, len(bl), n, np.ascontiguousarray(bl, dtype=np.int32),
                    np.ascontiguousarray(kl, dtype=np.int32), offs, start.astype(np.int32)))
                plan.counts[(cab, ccd, fp32)] = n
                if fp32:
                    plan.n32 += n
                else:
                    plan.n64 += n
    plan.n_screened = total - plan.n64 - plan.n32
    _launches_to_device(plan.launches, pt.dev)
    return plan


def _launches_to_device(launches, dev) -> None:
    """Replace the launches' host index arrays by contiguous slices of two
    device buffers (one int32, one int64): two host-to-device copies for the
    whole plan instead of four per launch."""
    if not launches:
        return
    i32 = [a for L in launches for a in (L.bra, L.ket, L.kstart)]
    n32 = [len(a) for a in i32]
    b32 = torch.from_numpy(np.concatenate(i32).astype(np.int32)).to(dev)
    b64 = torch.from_numpy(np.concatenate([L.offsets for L in launches])).to(dev)
    p32 = p64 = 0
    k = 0
    for L in launches:
        views = []
        for _ in range(3):
            views.append(b32[p32:p32 + n32[k]])
            p32 += n32[k]
            k += 1
        L.bra, L.ket, L.kstart = views
        m = len(L.offsets)
        L.offsets = b64[p64:p64 + m]
        p64 += m


def run_eri(pt: PairTable, plan: QuartetPlan, mode: int, out: torch.Tensor,
            only=None) -> None:
    """Launch every class of ``plan`` writing into ``out`` (dense n^4 tensor
    or tile-major supermatrix)."""
    nat.ensure_init(pt.dev)
    s = nat.stream_ptr()
    for L in plan.launches:
        if only is not None and (L.cab, L.ccd) not in only:
            continue
        a = pt.base_args(mode, out=out)
        a.bra_list = L.bra.data_ptr()
        a.ket_list = L.ket.data_ptr()
        a.offsets = L.offsets.data_ptr()
        a.kstart = L.kstart.data_ptr()
        a.nbra = L.nbra
        a.nquartets = L.nquartets
        a.variant = variant_for(L.cab, L.ccd)
        nat.call("qc_eri", L.cab, L.ccd, L.fp32, ctypes.byref(a), s)


# --- FLOP model (algorithmic work per primitive quartet) ------------------

def _nh(L):
    return (L + 1) * (L + 2) * (L + 3) // 6


def _nc(l):
    return (l + 1) * (l + 2) // 2


def flops_per_primitive_quartet(la, lb, lc, ld) -> int:
    """Boys + prefactor + R recursion + ket and bra Hermite contractions:
    (40+4L) + 30 + 3(L+1)(L+2)(L+3)(L+4)/24 + 2 N_H(Lab) N_H(Lcd) nc_c nc_d
    + 2 nc_a nc_b N_H(Lab) nc_c nc_d, bra = the pair with the larger L."""
    if la + lb < lc + ld:
        la, lb, lc, ld = lc, ld, la, lb
    L = la + lb + lc + ld
    Lab, Lcd = la + lb, lc + ld
    return ((40 + 4 * L) + 30 + 3 * (L + 1) * (L + 2) * (L + 3) * (L + 4) // 24
            + 2 * _nh(Lab) * _nh(Lcd) * _nc(lc) * _nc(ld)
            + 2 * _nc(la) * _nc(lb) * _nh(Lab) * _nc(lc) * _nc(ld))


def grad_flops_per_primitive_quartet(la, lb, lc, ld, ket_prims: float = 9.0) -> float:
    """Algorithmic FLOP of one primitive quartet of the derivative-ERI
    contraction (dA, dP = d(A+B), dC of sum G d(ab|cd)), in the form the
    kernels evaluate it:
    Boys (40 + 4(L+1)) + prefactor 30 + R to order L+1
    (3(L+2)(L+3)(L+4)(L+5)/24) + h for t+u+v <= Lab+1 (2 N_H(Lab+1) N_H(Lcd)
    nc_cd) + the three C-derivative lists for t+u+v <= Lab (6 N_H(Lab)
    N_H(Lcd+1) nc_cd) + their contraction with Y (2 nc_cd (3 N_H(Lab+1) +
    6 N_H(Lab))) + the bra-side Y = G E tables (8 nc_ab nc_cd N_H(Lab+1) per
    bra primitive, shared by the quartet's ket primitives: divided by
    ket_prims)."""
    if la + lb < lc + ld:
        la, lb, lc, ld = lc, ld, la, lb
    L = la + lb + lc + ld
    Lab, Lcd = la + lb, lc + ld
    ncd, nab = _nc(lc) * _nc(ld), _nc(la) * _nc(lb)
    return ((40 + 4 * (L + 1)) + 30 + 3 * (L + 2) * (L + 3) * (L + 4) * (L + 5) / 24
            + 2 * _nh(Lab + 1) * _nh(Lcd) * ncd + 6 * _nh(Lab) * _nh(Lcd + 1) * ncd
            + 2 * ncd * (3 * _nh(Lab + 1) + 6 * _nh(Lab))
            + 8 * nab * ncd * _nh(Lab + 1) / max(ket_prims, 1.0))


def plan_flops(pt: PairTable, plan: QuartetPlan, nprim_quartet: float | None = None) -> Dict[int, float]:
    """Algorithmic FLOPs of the plan split by precision (fp32 flag), using
    the mean primitive-pair count of each pair class."""
    mean_np = {}
    sp_n = pt.sp_nprim.cpu().numpy()
    for c in range(6):
        lst = pt.class_pairs[c]
        mean_np[c] = float(sp_n[lst].mean()) if len(lst) else 0.0
    out = {0: 0.0, 1: 0.0}
    for (cab, ccd, fp32), n in plan.counts.items():
        la, lb = CLASS_L[cab]
        lc, ld = CLASS_L[ccd]
        out[fp32] += n * mean_np[cab] * mean_np[ccd] * flops_per_primitive_quartet(la, lb, lc, ld)
    return out


def cart_to_sph_matrix(basis) -> np.ndarray:
    """Block-diagonal (n_sph x n_cart) matrix C with rows in the libcint
    spherical order (jax_qc/integrals/spherical.py:34-67): X_sph = C X_cart C^T
    and D_cart = C^T D_sph C."""
    r3 = 3.0 ** 0.5
    blocks = {0: np.eye(1), 1: np.eye(3),
              2: np.array([[0, r3, 0, 0, 0, 0], [0, 0, 0, 0, r3, 0], [-0.5, 0, 0, -0.5, 0, 1.0],
                           [0, 0, r3, 0, 0, 0], [0.5 * r3, 0, 0, -0.5 * r3, 0, 0]])}
    ls = [s.angular_momentum for s in basis.shells]
    ns = sum(2 * l + 1 for l in ls)
    nc = sum((l + 1) * (l + 2) // 2 for l in ls)
    C = np.zeros((ns, nc))
    i = j = 0
    for l in ls:
        b = blocks[l]
        C[i:i + b.shape[0], j:j + b.shape[1]] = b
        i += b.shape[0]
        j += b.shape[1]
    return C


def sm_tiles_per_dim(nbf: int) -> int:
    npair = nbf * (nbf + 1) // 2
    return (npair + SM_B - 1) // SM_B


def sm_numel(nbf: int) -> int:
    T = sm_tiles_per_dim(nbf)
    return T * (T + 1) // 2 * SM_B * SM_B
