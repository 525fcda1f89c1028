"""Batched RHF over molecules of uneven size.

Integrals: one pair table for all molecules (shell pairs never span two
molecules), one screened quartet plan whose every quartet has its bra and ket
pair in one molecule, one ERI launch per class pair writing each molecule's
J supermatrix into a shared pool (``qc_eri_args.sp_base``); P = J - 1/4 K is
assembled per molecule (Raffenetti, as ERIStore does for one molecule) and
optionally re-packed tile by tile in mixed precision.

SCF: the steps of jax_qc/scf/rhf.py:52-169 (core-Hamiltonian guess,
symmetric orthogonalisation, Pulay DIIS with the X^T(FDS - SDF)X error and a
ring of ``diis_space`` entries, convergence on |dE| and max|e|), batched over
molecules on matrices padded to the largest basis: S gets an identity block
and H a large diagonal shift in the padding, so every matrix stays block
diagonal, padded orbitals sort above every physical one, the padded density
is zero and the DIIS errors equal the unpadded ones.  A molecule is frozen
at the iteration it converges (its result is the one run_rhf returns).  The
Fock build of all molecules is pack -> tiles -> reduce -> unpack, captured
into one graph replayed once per iteration.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from jax_qc import _native as nat
from jax_qc.basis import build_basis_set
from jax_qc.integrals import engine as eng
from jax_qc.integrals.one_electron import nuclear_repulsion_energy
from jax_qc.integrals.store import TILE_F32, TILE_F64, TILE_ZERO

PAD_SHIFT = 1.0e4  # Eh on the padded diagonal of H (above any orbital energy here)


class BatchPairTable:
    """Concatenation of per-molecule pair tables (same field names as
    ``engine.PairTable``; ``sp_ao`` stays molecule-local) plus the molecule
    of each shell pair and the element offset of its molecule's
    supermatrix in the J/P pools."""

    def __init__(self, tables: List[eng.PairTable], dev=None):
        self.dev = dev or tables[0].dev
        self.tables = tables
        self.nmol = len(tables)
        self.nbf_m = np.array([t.nbf for t in tables], dtype=np.int64)
        self.nbf = int(self.nbf_m.max())
        self.T_m = np.array([eng.sm_tiles_per_dim(int(n)) for n in self.nbf_m], dtype=np.int64)
        numel = self.T_m * (self.T_m + 1) // 2 * eng.SM_B * eng.SM_B
        self.sm_base_m = np.zeros(self.nmol + 1, dtype=np.int64)
        self.sm_base_m[1:] = np.cumsum(numel)
        npairs = np.array([t.npairs for t in tables], dtype=np.int64)
        nprim = np.array([t.nprim_pairs for t in tables], dtype=np.int64)
        pair0 = np.concatenate([[0], np.cumsum(npairs)[:-1]])
        prim0 = np.concatenate([[0], np.cumsum(nprim)[:-1]])
        self.npairs = int(npairs.sum())
        self.nprim_pairs = int(nprim.sum())
        self.sp_mol = np.repeat(np.arange(self.nmol, dtype=np.int32), npairs)
        self.sp_class = np.concatenate([t.sp_class for t in tables])
        cat = lambda name: torch.cat([getattr(t, name) for t in tables])
        self.pairs64 = cat("pairs64")
        self.pairs32 = torch.cat([t.pairs32 for t in tables]) if tables[0].pairs32 is not None else None
        self.sp_prim_off = torch.cat([t.sp_prim_off + int(p0) for t, p0 in zip(tables, prim0)])
        self.sp_nprim = cat("sp_nprim")
        self.sp_ao = cat("sp_ao")
        self.sp_base = torch.from_numpy(self.sm_base_m[self.sp_mol]).to(self.dev)
        self.class_pairs: Dict[int, np.ndarray] = {
            c: np.nonzero(self.sp_class == c)[0].astype(np.int32) for c in range(6)}
        self.Q = None

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
        a.sp_base = self.sp_base.data_ptr() if mode == eng.MODE_SUPER else 0
        return a


def plan_quartets_batched(bpt: BatchPairTable, tau: float = 1e-12,
                          tau32: Optional[float] = None) -> eng.QuartetPlan:
    """``engine.plan_quartets`` with quartets restricted to one molecule:
    every class list is sorted by (molecule, -Q), so a bra pair's surviving
    kets are a prefix of its own molecule's segment of the ket list."""
    if bpt.Q is None:
        eng.schwarz(bpt)
    Q, mol = bpt.Q, bpt.sp_mol
    plan = eng.QuartetPlan(tau=tau, tau32=tau32 or 0.0)
    lists, qs, segs = {}, {}, {}
    for c in range(6):
        lst = bpt.class_pairs[c]
        order = np.lexsort((-Q[lst], mol[lst]))
        lists[c] = lst[order]
        qs[c] = Q[lst][order]
        m = mol[lists[c]]
        segs[c] = np.searchsorted(m, np.arange(bpt.nmol + 1), side="left")
    total = 0
    for cab in range(6):
        for ccd in range(cab + 1):
            bl, kl = lists[cab], lists[ccd]
            if len(bl) == 0 or len(kl) == 0:
                continue
            qb, qk = qs[cab], qs[ccd]
            kstart = np.zeros(len(bl), dtype=np.int64)
            kmax = np.zeros(len(bl), dtype=np.int64)
            k64 = np.zeros(len(bl), dtype=np.int64)
            sb, sk = segs[cab], segs[ccd]
            for m in range(bpt.nmol):
                b0, b1, c0, c1 = sb[m], sb[m + 1], sk[m], sk[m + 1]
                if b1 == b0 or c1 == c0:
                    continue
                neg = -qk[c0:c1]
                qbm = np.maximum(qb[b0:b1], 1e-300)
                with np.errstate(divide="ignore"):
                    hi = np.searchsorted(neg, -(tau / qbm), side="right")
                    lo64 = (np.searchsorted(neg, -(tau32 / qbm), side="right") if tau32 else hi)
                nb, nk = b1 - b0, c1 - c0
                total += nb * nk if cab != ccd else nb * (nb + 1) // 2
                if cab == ccd:
                    cap = np.arange(1, nb + 1)
                    hi = np.minimum(hi, cap)
                    lo64 = np.minimum(lo64, cap)
                kstart[b0:b1] = c0
                kmax[b0:b1] = hi
                k64[b0:b1] = np.minimum(lo64, hi)
            for fp32, (start, stop) in ((0, (np.zeros_like(k64), k64)), (1, (k64, kmax))):
                cnt = (stop - start).astype(np.int64)
                n = int(cnt.sum())
                if n == 0:
                    continue
                offs = np.zeros(len(bl) + 1, dtype=np.int64)
                offs[1:] = np.cumsum(cnt)
                d = bpt.dev
                plan.launches.append(eng.ClassLaunch(
                    cab, ccd, fp32, len(bl), n,
                    torch.from_numpy(np.ascontiguousarray(bl)).to(d),
                    torch.from_numpy(np.ascontiguousarray(kl)).to(d),
                    torch.from_numpy(offs).to(d),
                    torch.from_numpy((kstart + start).astype(np.int32)).to(d)))
                plan.counts[(cab, ccd, fp32)] = n
                if fp32:
                    plan.n32 += n
                else:
                    plan.n64 += n
    plan.n_screened = total - plan.n64 - plan.n32
    return plan


@dataclass
class BatchResult:
    energies: np.ndarray        # total energies (Eh)
    converged: np.ndarray       # bool per molecule
    n_iterations: np.ndarray    # per molecule
    iterations: int             # batched iterations run
    fock_seconds: float         # summed wall time of the batched Fock builds
    scf_seconds: float
    orbital_energies: Optional[torch.Tensor] = None
    density: Optional[torch.Tensor] = None


class BatchedRHF:
    """Integrals and RHF for a list of closed-shell molecules.

    tau: Schwarz threshold; tau32: quartet bound below which ERIs are summed
    in FP32 (None: all FP64); tile_f32: supermatrix tiles whose max element
    is below it are stored in FP32 (None: all FP64); seg: tiles per Fock
    work item; eri_tuning: {(cab, ccd): (variant, block)} for the ERI
    launches (from ``jax_qc.autotune.autotune_batch``)."""

    def __init__(self, mols, basis="even-tempered", tau: float = 1e-12,
                 tau32: Optional[float] = None, tile_f32: Optional[float] = None,
                 seg: int = 4, eri_tuning: Optional[dict] = None, dev=None):
        self.dev = dev or nat.device()
        nat.ensure_init(self.dev)
        self.mols = list(mols)
        self.nmol = len(self.mols)
        for m in self.mols:
            if m.spin != 0 or m.n_electrons % 2:
                raise ValueError("BatchedRHF needs closed-shell molecules (spin 0, even electrons)")
        t0 = time.perf_counter()
        self.bases = [build_basis_set(m, basis) if isinstance(basis, str) else basis[i]
                      for i, m in enumerate(self.mols)]
        tables = [eng.PairTable(b, self.dev) for b in self.bases]
        self.bpt = BatchPairTable(tables, self.dev)
        self.nbf_m = self.bpt.nbf_m
        self.nmax = self.bpt.nbf
        self.n_occ = np.array([m.n_electrons // 2 for m in self.mols], dtype=np.int64)
        self.E_nuc = np.array([nuclear_repulsion_energy(m) for m in self.mols])
        n, N = self.nmol, self.nmax
        S, T, V = self._one_electron()
        pad = torch.zeros(n, N, dtype=torch.float64, device=self.dev)
        for i in range(n):
            pad[i, int(self.nbf_m[i]):] = 1.0
        self.S = S + torch.diag_embed(pad)
        self.H = T + V + PAD_SHIFT * torch.diag_embed(pad)
        self.one_electron_seconds = time.perf_counter() - t0

        # ---- ERIs: one launch per class pair for every molecule ----------
        t0 = time.perf_counter()
        self.tau, self.tau32 = tau, tau32
        self.plan = plan_quartets_batched(self.bpt, tau, tau32)
        J = torch.zeros(int(self.bpt.sm_base_m[-1]), dtype=torch.float64, device=self.dev)
        self.eri_tuning = dict(eri_tuning or {})
        for L in self.plan.launches:
            v, blk = self.eri_tuning.get((L.cab, L.ccd), (eng.variant_for(L.cab, L.ccd), 0))
            self.eri_launch(L, J, v, blk)
        torch.cuda.synchronize(self.dev)
        self.eri_seconds = time.perf_counter() - t0
        # ---- P = J - 1/4 K per molecule ------------------------------------
        P = torch.empty_like(J)
        base = self.bpt.sm_base_m
        s = nat.stream_ptr()
        for i in range(n):
            nat.call("qc_sm_assemble", 0, J.data_ptr() + 8 * int(base[i]),
                     P.data_ptr() + 8 * int(base[i]), int(self.nbf_m[i]), s)
        del J
        self.ntiles = int(base[-1]) // (eng.SM_B * eng.SM_B)
        self.P, self.fmt, self.off = P, None, None
        self.tile_counts = {"f64": self.ntiles, "f32": 0, "zero": 0}
        if tile_f32 is not None:
            self._compact(tile_f32)
        torch.cuda.synchronize(self.dev)
        self.assembly_seconds = time.perf_counter() - t0 - self.eri_seconds
        self._setup_fock(seg)
        self.graph = None

    def _one_electron(self):
        """S, T, V of every molecule, padded to nmax: one launch per pair
        class and matrix for all molecules (qc_one_electron_batch)."""
        bpt, dev, N = self.bpt, self.dev, self.nmax
        nat_m = np.array([len(m.atomic_numbers) for m in self.mols], dtype=np.int64)
        at0 = np.concatenate([[0], np.cumsum(nat_m)[:-1]])
        xyz = torch.from_numpy(np.ascontiguousarray(
            np.concatenate([np.asarray(m.coords, dtype=np.float64) for m in self.mols]))).to(dev)
        zz = torch.from_numpy(np.concatenate(
            [np.asarray(m.atomic_numbers, dtype=np.float64) for m in self.mols])).to(dev)
        mol = bpt.sp_mol
        pair_atoms = torch.from_numpy(
            np.stack([at0[mol], nat_m[mol]], 1).astype(np.int32).reshape(-1)).to(dev)
        pair_base = torch.from_numpy(mol.astype(np.int64) * N * N).to(dev)
        out = []
        for which in (0, 1, 2):
            M = torch.zeros(self.nmol, N, N, dtype=torch.float64, device=dev)
            for c in range(6):
                lst = bpt.class_pairs[c]
                if len(lst) == 0:
                    continue
                ld = torch.from_numpy(lst).to(dev)
                nat.call("qc_one_electron_batch", which, c, nat.ptr(bpt.pairs64),
                         nat.ptr(bpt.sp_prim_off), nat.ptr(bpt.sp_nprim), nat.ptr(bpt.sp_ao),
                         nat.ptr(ld), len(lst), nat.ptr(pair_atoms), nat.ptr(xyz), nat.ptr(zz), N,
                         nat.ptr(pair_base), nat.ptr(M), nat.stream_ptr())
            out.append(M)
        return out

    def eri_launch(self, L: eng.ClassLaunch, J: torch.Tensor, variant: int, block: int = 0):
        """One class launch of the batch plan into the J pool ``J``."""
        a = self.bpt.base_args(eng.MODE_SUPER, out=J)
        a.bra_list = L.bra.data_ptr()
        a.ket_list = L.ket.data_ptr()
        a.offsets = L.offsets.data_ptr()
        a.kstart = L.kstart.data_ptr()
        a.nbra = L.nbra
        a.nquartets = L.nquartets
        a.variant = variant
        a.block = block
        nat.call("qc_eri", L.cab, L.ccd, L.fp32, ctypes.byref(a), nat.stream_ptr())

    # -- mixed-precision tiles ----------------------------------------------
    def _compact(self, tau_f32: float, tau_zero: float = 1e-14):
        tmax = torch.empty(self.ntiles, dtype=torch.float64, device=self.dev)
        nat.call("qc_tiles_maxabs", nat.ptr(self.P), self.ntiles, nat.ptr(tmax), nat.stream_ptr())
        tm = tmax.cpu().numpy()
        fmt = np.where(tm < tau_zero, TILE_ZERO, np.where(tm < tau_f32, TILE_F32, TILE_F64))
        fmt = fmt.astype(np.uint8)
        size = np.where(fmt == TILE_F64, 8 * 4096, np.where(fmt == TILE_F32, 4 * 4096, 0))
        off = np.zeros(self.ntiles, dtype=np.int64)
        off[1:] = np.cumsum(size)[:-1]
        total = int(size.sum())
        pool = torch.empty(max(16, total), dtype=torch.uint8, device=self.dev)
        self.fmt = torch.from_numpy(fmt).to(self.dev)
        self.off = torch.from_numpy(off).to(self.dev)
        nat.call("qc_tiles_pack", nat.ptr(self.P), self.ntiles, nat.ptr(self.fmt), nat.ptr(self.off),
                 nat.ptr(pool), nat.stream_ptr())
        torch.cuda.synchronize(self.dev)
        self.P = pool
        self.tile_counts = {"f64": int((fmt == TILE_F64).sum()), "f32": int((fmt == TILE_F32).sum()),
                            "zero": int((fmt == TILE_ZERO).sum())}

    @property
    def matrix_bytes(self) -> int:
        """Bytes of supermatrix one batched Fock build streams."""
        return int(self.P.numel() * self.P.element_size())

    # -- the batched Fock build ---------------------------------------------
    def _setup_fock(self, seg: int):
        T = self.bpt.T_m
        self.seg = int(seg)
        self.molT = torch.from_numpy(T.copy()).to(self.dev)
        tile0 = np.zeros(self.nmol + 1, dtype=np.int64)
        tile0[1:] = np.cumsum(T * (T + 1) // 2)
        row0 = np.zeros(self.nmol + 1, dtype=np.int64)
        row0[1:] = np.cumsum(T)
        item0 = np.zeros(self.nmol + 1, dtype=np.int64)
        Tc = np.ascontiguousarray(T, dtype=np.int64)
        self.nitems = int(nat.lib().qc_batch_items(
            Tc.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), self.nmol, self.seg,
            item0.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self.tile0 = torch.from_numpy(tile0).to(self.dev)
        self.row0 = torch.from_numpy(row0).to(self.dev)
        self.item0 = torch.from_numpy(item0).to(self.dev)
        self.nbf_dev = torch.from_numpy(self.nbf_m.copy()).to(self.dev)
        self.nrows = int(row0[-1])
        self.npad = self.nrows * eng.SM_B
        self.x = torch.empty(self.npad, dtype=torch.float64, device=self.dev)
        self.y = torch.empty(self.npad, dtype=torch.float64, device=self.dev)
        self.work = torch.empty((self.nitems + self.ntiles) * eng.SM_B, dtype=torch.float64,
                                device=self.dev)
        self.D_static = torch.zeros(self.nmol, self.nmax, self.nmax, dtype=torch.float64,
                                    device=self.dev)
        self.F_static = torch.empty_like(self.D_static)

    def set_seg(self, seg: int):
        """Change the tiles per Fock work item (drops a captured graph)."""
        self._setup_fock(seg)
        self.graph = None

    def fock_into(self, D: torch.Tensor, F: torch.Tensor):
        """F = H + P D~ for every molecule (four launches, no allocation)."""
        s = nat.stream_ptr()
        nat.call("qc_batch_pack_density", nat.ptr(D), self.nmax, self.nmol, nat.ptr(self.nbf_dev),
                 nat.ptr(self.row0), 1, nat.ptr(self.x), self.npad, s)
        nat.call("qc_batch_symv_tiles", nat.ptr(self.P), nat.ptr(self.fmt), nat.ptr(self.off),
                 self.nmol, nat.ptr(self.molT), nat.ptr(self.tile0), nat.ptr(self.item0),
                 nat.ptr(self.row0), self.nitems, self.ntiles, self.npad, 1, self.seg,
                 nat.ptr(self.x), nat.ptr(self.work), s)
        nat.call("qc_batch_symv_reduce", self.nmol, nat.ptr(self.molT), nat.ptr(self.tile0),
                 nat.ptr(self.item0), nat.ptr(self.row0), self.nrows, self.nitems, self.npad, 1,
                 self.seg, nat.ptr(self.y), nat.ptr(self.work), s)
        nat.call("qc_batch_unpack_fock", nat.ptr(self.H), nat.ptr(self.y), 1.0, self.nmax,
                 self.nmol, nat.ptr(self.nbf_dev), nat.ptr(self.row0), 1, self.npad, nat.ptr(F), s)

    def capture(self):
        """Capture the batched Fock build of D_static into F_static as one graph."""
        side = torch.cuda.Stream(self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            self.fock_into(self.D_static, self.F_static)
        torch.cuda.current_stream(self.dev).wait_stream(side)
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.fock_into(self.D_static, self.F_static)
        torch.cuda.synchronize(self.dev)
        self.graph = g
        return g

    def fock(self, D: torch.Tensor, use_graph: bool = True) -> torch.Tensor:
        if use_graph:
            if self.graph is None:
                self.capture()
            self.D_static.copy_(D)
            self.graph.replay()
            return self.F_static.clone()
        F = torch.empty_like(D)
        self.fock_into(D.contiguous(), F)
        return F

    # -- SCF ------------------------------------------------------------------
    def _diag(self, F, X):
        e, Cp = torch.linalg.eigh(X @ F @ X)
        return e, X @ Cp

    def _density(self, C):
        mask = (torch.arange(self.nmax, device=self.dev)[None, :] <
                torch.as_tensor(self.n_occ, device=self.dev)[:, None]).to(C.dtype)
        Co = C * mask[:, None, :]
        return 2.0 * Co @ Co.transpose(1, 2)

    def _energy(self, D, F):
        return 0.5 * (D * (self.H + F)).sum(dim=(1, 2))

    def run(self, conv: float = 1e-10, max_iter: int = 100, diis_space: int = 8,
            use_graph: bool = True) -> BatchResult:
        t_start = time.perf_counter()
        fock_t = 0.0
        w, U = torch.linalg.eigh(self.S)
        X = (U * (1.0 / torch.sqrt(w))[:, None, :]) @ U.transpose(1, 2)
        X = 0.5 * (X + X.transpose(1, 2))
        _, C = self._diag(self.H, X)
        D = self._density(C)
        t = time.perf_counter()
        F0 = self.fock(D, use_graph)
        torch.cuda.synchronize(self.dev)
        fock_t += time.perf_counter() - t
        orb_e, C = self._diag(F0, X)
        E = self._energy(D, F0)
        last = E.cpu().numpy()
        focks, errs = [], []
        # <e_i, e_j> per molecule, kept on the device in step with the ring
        gram = torch.zeros(self.nmol, 0, 0, dtype=torch.float64, device=self.dev)
        done = np.zeros(self.nmol, dtype=bool)
        n_it = np.zeros(self.nmol, dtype=np.int64)
        E_final = last.copy()
        it = 0
        for it in range(1, max_iter + 1):
            t = time.perf_counter()
            F = self.fock(D, use_graph)
            torch.cuda.synchronize(self.dev)
            fock_t += time.perf_counter() - t
            err = X @ (F @ D @ self.S - self.S @ D @ F) @ X
            if len(focks) == diis_space:
                focks.pop(0)
                errs.pop(0)
                gram = gram[:, 1:, 1:]
            focks.append(F)
            errs.append(err)
            row = torch.stack([(x * err).sum(dim=(1, 2)) for x in errs], 1)
            m = len(errs)
            G = torch.empty(self.nmol, m, m, dtype=torch.float64, device=self.dev)
            G[:, :-1, :-1] = gram
            G[:, -1, :] = row
            G[:, :, -1] = row
            gram = G
            if m >= 2:
                # Pulay system per molecule, all molecules in one batched solve;
                # an exactly singular system keeps the un-extrapolated Fock
                # (jax_qc/scf/diis.py:97-100)
                B = torch.zeros(self.nmol, m + 1, m + 1, dtype=torch.float64, device=self.dev)
                B[:, :m, :m] = G
                B[:, m, :m] = -1.0
                B[:, :m, m] = -1.0
                rhs = torch.zeros(self.nmol, m + 1, dtype=torch.float64, device=self.dev)
                rhs[:, m] = -1.0
                sol, info = torch.linalg.solve_ex(B, rhs)
                fallback = torch.zeros(self.nmol, m, dtype=torch.float64, device=self.dev)
                fallback[:, -1] = 1.0
                ct = torch.where((info == 0)[:, None], sol[:, :m], fallback)
                Fx = torch.einsum("ni,nijk->njk", ct, torch.stack(focks, 1))
            else:
                Fx = F
            orb_new, C_new = self._diag(Fx, X)
            D_new = self._density(C_new)
            E_new = self._energy(D_new, Fx)
            En = E_new.cpu().numpy()
            emax = err.abs().amax(dim=(1, 2)).cpu().numpy()
            active = ~done
            n_it[active] = it
            conv_now = active & (np.abs(En - last) < conv) & (emax < max(conv, 1e-6))
            keep = torch.from_numpy(active).to(self.dev)[:, None, None]
            D = torch.where(keep, D_new, D)
            orb_e = torch.where(keep[:, :, 0], orb_new, orb_e)
            E_final = np.where(active, En, E_final)
            last = np.where(active, En, last)
            done |= conv_now
            if done.all():
                break
        torch.cuda.synchronize(self.dev)
        return BatchResult(E_final + self.E_nuc, done.copy(), n_it, it, fock_t,
                           time.perf_counter() - t_start, orb_e, D)
