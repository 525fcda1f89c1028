"""ERIStore: the unique ERIs of one geometry held in HBM as packed
supermatrices, and the Fock builds that stream them.

Layout: rows/columns are basis-function pairs p = i(i+1)/2 + j (i >= j);
64 x 64 tiles (R, C), C <= R, tile-major (C2's 240 functions: 3.35 GB FP64
per supermatrix; 480 functions: 53 GB).  Two supermatrices can be held:

* ``P`` (RHF): P[(ij),(kl)] = (ij|kl) - 1/4[(ik|jl) + (il|jk)], so
  F = H + P D~ with D~_kl = D_kl (2 - delta_kl) -- one symmetric
  matrix-vector product per Fock build (Raffenetti 1973).
* ``J``, ``K`` (UHF / separate J and K): J[(ij),(kl)] = (ij|kl),
  K[(ij),(kl)] = 1/2[(ik|jl) + (il|jk)].

This is the integral-storage counterpart of the reference's dense
``IntegralSet.ERI`` (jax_qc/integrals/eri.py:69-105), consumed by
``build_fock_rhf/uhf`` exactly where the reference contracts its dense tensor
(jax_qc/scf/fock.py:31-79).
"""

from __future__ import annotations

import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from jax_qc import _native as nat
from jax_qc.integrals import engine as eng


TILE_ZERO, TILE_F32, TILE_F64, TILE_I16, TILE_I8 = 0, 1, 2, 3, 4
# pool bytes per tile by format: 64 x 64 elements; the integer formats carry
# a 16-byte header (the tile's FP64 scale, then padding)
TILE_BYTES = {TILE_ZERO: 0, TILE_F32: 4 * 4096, TILE_F64: 8 * 4096,
              TILE_I16: 16 + 2 * 4096, TILE_I8: 16 + 4096}


# FP64 supermatrix buffers kept for the next store of the same size (J and
# the K or P assembled beside it).  Back-to-back evaluations at one basis size
# (a geometry optimisation) then allocate no new 10-100 GB blocks: torch's
# allocator leaves small long-lived tensors inside a freed supermatrix's
# segment, so the next store's J needed a fresh device allocation, which
# took 0.4-2 s per (H2O)_16 evaluation.  A buffer is free when nothing but
# this list refers to it (views do not count: a caller that keeps a view of a
# store's FP64 supermatrix past the store's life must copy it instead);
# release_supermatrix_buffers() hands the memory back.
_SM_BUFFERS: list = []
SM_BUFFERS_MAX = 2


def _sm_buffer(numel: int, dev: torch.device, zero: bool) -> torch.Tensor:
    """A float64 vector of numel elements on dev, from the kept buffers when
    one of that size is free (zeroed if asked), else newly allocated (and
    kept while fewer than SM_BUFFERS_MAX are)."""
    for i in range(len(_SM_BUFFERS)):
        # a free buffer has two references: the list and getrefcount's argument
        if (sys.getrefcount(_SM_BUFFERS[i]) == 2 and _SM_BUFFERS[i].numel() == numel
                and _SM_BUFFERS[i].device == dev):
            t = _SM_BUFFERS[i]
            return t.zero_() if zero else t
    for i in reversed(range(len(_SM_BUFFERS))):  # free buffers of other sizes go
        if sys.getrefcount(_SM_BUFFERS[i]) == 2 and _SM_BUFFERS[i].numel() != numel:
            del _SM_BUFFERS[i]
    if zero:
        t = torch.zeros(numel, dtype=torch.float64, device=dev)
    else:
        t = torch.empty(numel, dtype=torch.float64, device=dev)
    if len(_SM_BUFFERS) < SM_BUFFERS_MAX:
        _SM_BUFFERS.append(t)
    return t


def release_supermatrix_buffers() -> None:
    """Drop the kept FP64 supermatrix buffers (those no store still uses
    are freed) and return torch's cached device memory."""
    _SM_BUFFERS.clear()
    if torch.cuda.is_available():
        torch.cuda.empty_cache()


@dataclass
class CompactSupermatrix:
    """A supermatrix re-packed tile by tile: tiles whose max |element| m is
    below ``tau_zero`` are dropped; below ``tau_i8`` / ``tau_i16`` they are
    held as int8 / int16 multiples of a per-tile scale m/127 or m/32767 (each
    element off by at most m/254 or m/65534); below ``tau_f32`` in FP32; the
    rest in FP64.  The Fock build accumulates in FP64 either way."""

    pool: torch.Tensor      # uint8 byte pool
    fmt: torch.Tensor       # uint8 per tile
    off: torch.Tensor       # int64 byte offset per tile
    tmax: torch.Tensor      # float32 max |element| per tile (tile screening)
    n_zero: int
    n_f32: int
    n_f64: int
    bytes: int
    tau_f32: float
    tau_zero: float
    n_i16: int = 0
    n_i8: int = 0
    tau_i16: float = 0.0
    tau_i8: float = 0.0


def tile_formats(tm: np.ndarray, tau_f32: float, tau_zero: float = 1e-14,
                 tau_i16: float = 0.0, tau_i8: float = 0.0) -> np.ndarray:
    """Per-tile format codes from the tiles' max |element| tm (a NumPy
    array or a device tensor; the result is of the same kind): zero below
    tau_zero, int8 below tau_i8, int16 below tau_i16, FP32 below tau_f32,
    FP64 otherwise."""
    if isinstance(tm, torch.Tensor):
        fmt = torch.full(tm.shape, TILE_F64, dtype=torch.uint8, device=tm.device)
    else:
        fmt = np.full(tm.shape, TILE_F64, dtype=np.uint8)
    fmt[tm < tau_f32] = TILE_F32
    fmt[tm < tau_i16] = TILE_I16
    fmt[tm < tau_i8] = TILE_I8
    fmt[tm < tau_zero] = TILE_ZERO
    return fmt


class ERIStore:
    def __init__(self, basis, tau: float = 1e-12, tau32: float | None = None,
                 dev=None, pair_table=None):
        """Compute the screened unique ERIs of ``basis`` into the J
        supermatrix; P (RHF) and K are assembled from it on first use.
        tau: Schwarz threshold; tau32: bound below which quartets are summed
        in FP32 (None = all FP64)."""
        self.dev = dev or nat.device()
        nat.ensure_init(self.dev)
        self.basis = basis
        self.nbf = int(basis.n_basis)
        self.tau, self.tau32 = tau, tau32
        self.pt = pair_table or eng.PairTable(basis, self.dev)
        self.plan = eng.plan_quartets(self.pt, tau, tau32)
        self.T = eng.sm_tiles_per_dim(self.nbf)
        self.npad = self.T * eng.SM_B
        t0 = time.perf_counter()
        self.J = _sm_buffer(eng.sm_numel(self.nbf), self.dev, zero=True)
        eng.run_eri(self.pt, self.plan, eng.MODE_SUPER, self.J)
        torch.cuda.synchronize(self.dev)
        self.eri_seconds = time.perf_counter() - t0
        self.P = self.K = None
        # per-tile max |element| of P and K, written by the kernels that
        # assemble them (what compact() needs, without another read of the
        # matrix); dropped once used
        self._tmax = {}
        self._work, self._x, self._y, self._xm = {}, {}, {}, {}
        self.inc = None

    def use_compact(self, tau_f32: float = 1e-4, tau_zero: float = 1e-14,
                    tau_i16: float = 0.0, tau_i8: float = 0.0):
        """Replace the assembled supermatrices (P for RHF; J and K for UHF)
        by their compacted mixed forms.  With neither P nor K assembled yet,
        P is assembled first (the RHF case); a store that already holds K is
        taken as UHF and no P is built."""
        kw = dict(tau_f32=tau_f32, tau_zero=tau_zero, tau_i16=tau_i16, tau_i8=tau_i8)
        if self.P is None and self.K is None:
            self.ensure_P()
        if isinstance(self.P, torch.Tensor):
            self.P = self.compact(self.P, tmax=self._tmax.pop("P", None), **kw)
        if isinstance(self.K, torch.Tensor):
            self.K = self.compact(self.K, tmax=self._tmax.pop("K", None), **kw)
        if isinstance(self.J, torch.Tensor) and self.K is not None:
            self.J = self.compact(self.J, **kw)
        return self

    def _assemble_from_J(self, what):
        if not isinstance(self.J, torch.Tensor):
            raise RuntimeError(
                f"cannot assemble {what}: the J supermatrix was "
                + ("released" if self.J is None else "compacted") +
                "; assemble P or K before use_compact() / release_J()")

    def ensure_P(self):
        if self.P is None:
            self._assemble_from_J("P")
            self.P = _sm_buffer(self.J.numel(), self.dev, zero=False)
            self._tmax["P"] = self._tile_vector()
            nat.call("qc_sm_assemble_maxabs", 0, nat.ptr(self.J), nat.ptr(self.P), self.nbf,
                     nat.ptr(self._tmax["P"]), nat.stream_ptr())
        return self.P

    def _tile_vector(self):
        return torch.empty(self.T * (self.T + 1) // 2, dtype=torch.float64, device=self.dev)

    def ensure_K(self):
        if self.K is None:
            self._assemble_from_J("K")
            self.K = _sm_buffer(self.J.numel(), self.dev, zero=False)
            self._tmax["K"] = self._tile_vector()
            nat.call("qc_sm_assemble_maxabs", 1, nat.ptr(self.J), nat.ptr(self.K), self.nbf,
                     nat.ptr(self._tmax["K"]), nat.stream_ptr())
        return self.K

    def assemble_p_k(self):
        """K from J, then J turned into P = J - K/2 in place: the pair of
        supermatrices the UHF Fock builds use (see fock_uhf), without a third
        n_pair^2 buffer or a second gather pass.  Diagonal tiles hold both
        triangles and the padding is zero in J and K alike, so this equals
        the assembly of P from J element by element."""
        K = self.ensure_K()
        if self.P is None:
            self._assemble_from_J("P")
            self._tmax["P"] = self._tile_vector()
            nat.call("qc_sm_p_from_jk", nat.ptr(self.J), nat.ptr(K), self.nbf,
                     nat.ptr(self._tmax["P"]), nat.stream_ptr())
            self.P, self.J = self.J, None
        return self.P, K

    def release_J(self):
        """Drop the J supermatrix once only P is needed (RHF)."""
        self.ensure_P()
        self.J = None

    def compact(self, M: torch.Tensor, tau_f32: float = 1e-4, tau_zero: float = 1e-14,
                tau_i16: float = 0.0, tau_i8: float = 0.0,
                tmax: torch.Tensor | None = None) -> CompactSupermatrix:
        """Mixed-precision, zero-skipping copy of supermatrix M (formats by
        tile max |element|: see CompactSupermatrix; tau_i16 = tau_i8 = 0 keeps
        the integer formats off).  tmax: M's per-tile max |element| if the
        caller has it (from the kernel that assembled M), else computed."""
        if tau_i8 > max(tau_i16, tau_f32) or tau_i16 > tau_f32:
            raise ValueError("tile thresholds must satisfy tau_i8 <= tau_i16 <= tau_f32")
        T = self.T
        nt = T * (T + 1) // 2
        if tmax is None:
            tmax = torch.empty(nt, dtype=torch.float64, device=self.dev)
            nat.call("qc_sm_tile_maxabs", nat.ptr(M), self.nbf, nat.ptr(tmax), nat.stream_ptr())
        elif tmax.shape != (nt,) or tmax.dtype != torch.float64:
            raise ValueError(f"tmax must be float64 of shape ({nt},)")
        # formats, sizes and pool offsets on the device; only the five format
        # counts come back to the host (to size the pool)
        fd = tile_formats(tmax, tau_f32, tau_zero, tau_i16, tau_i8)
        nbytes = torch.tensor([TILE_BYTES[f] for f in range(len(TILE_BYTES))], dtype=torch.int64,
                              device=self.dev)
        size = nbytes[fd.long()]
        od = (torch.cumsum(size, 0) - size).contiguous()
        counts = torch.bincount(fd.long(), minlength=len(TILE_BYTES)).cpu().tolist()
        cnt = {f: int(counts[f]) for f in TILE_BYTES}
        total = sum(cnt[f] * TILE_BYTES[f] for f in TILE_BYTES)
        pool = torch.empty(max(16, total), dtype=torch.uint8, device=self.dev)
        nat.call("qc_sm_pack", nat.ptr(M), self.nbf, nat.ptr(fd), nat.ptr(od), nat.ptr(pool),
                 nat.stream_ptr())
        # FP32 maxima rounded up, so a screen never undershoots a tile's bound
        t32 = torch.nextafter(tmax.float(), torch.full_like(tmax, float("inf"), dtype=torch.float32))
        return CompactSupermatrix(pool, fd, od, t32, cnt[TILE_ZERO], cnt[TILE_F32], cnt[TILE_F64],
                                  total, tau_f32, tau_zero, cnt[TILE_I16], cnt[TILE_I8],
                                  tau_i16, tau_i8)

    # -- Fock builds --------------------------------------------------------
    def bytes_per_matrix(self) -> int:
        return eng.sm_numel(self.nbf) * 8

    def _buffers(self, nrhs):
        if nrhs not in self._work:
            wb = nat.lib().qc_symv_workspace_bytes(self.nbf, nrhs)
            self._work[nrhs] = torch.empty(max(1, wb // 8), dtype=torch.float64, device=self.dev)
            self._x[nrhs] = torch.empty(nrhs * self.npad, dtype=torch.float64, device=self.dev)
            self._y[nrhs] = torch.empty(nrhs * self.npad, dtype=torch.float64, device=self.dev)
            self._xm[nrhs] = torch.empty(self.T, dtype=torch.float32, device=self.dev)
        return self._work[nrhs], self._x[nrhs], self._y[nrhs]

    def contract(self, M: torch.Tensor, D: torch.Tensor, H: torch.Tensor | None = None,
                 scale: float = 1.0, tau_skip: float = 0.0) -> torch.Tensor:
        """Return H + scale * M D~ for a density or a stack of densities.
        For a compacted M, tiles with max|M_tile| max|D~ block| < tau_skip
        are left out (each moves an element by less than 64 tau_skip); zero
        tiles are always left out."""
        D = D.contiguous()
        squeeze = D.dim() == 2
        if squeeze:
            D = D.unsqueeze(0)
        nrhs = D.shape[0]
        work, x, y = self._buffers(nrhs)
        s = nat.stream_ptr()
        nat.call("qc_pack_density", nat.ptr(D), self.nbf, nrhs, nat.ptr(x), s)
        if isinstance(M, CompactSupermatrix):
            xm = self._xm[nrhs]
            tau = float(tau_skip)
            nat.call("qc_block_maxabs", nat.ptr(x), self.nbf, nrhs, nat.ptr(xm), s)
            nat.call("qc_symv_screened_tiles", nat.ptr(M.pool), nat.ptr(M.fmt), nat.ptr(M.off),
                     nat.ptr(M.tmax), self.nbf, nrhs, nat.ptr(x), nat.ptr(xm), tau,
                     nat.ptr(work), s)
            nat.call("qc_symv_screened_reduce", self.nbf, nrhs, nat.ptr(y), nat.ptr(work),
                     nat.ptr(M.fmt), nat.ptr(M.tmax), nat.ptr(xm), tau, s)
        elif tau_skip > 0.0:
            raise ValueError("tile screening needs a compacted supermatrix (use_compact)")
        else:
            nat.call("qc_symv", nat.ptr(M), self.nbf, nrhs, nat.ptr(x), nat.ptr(y), nat.ptr(work), s)
        F = torch.empty_like(D)
        Hp = H.contiguous() if H is not None else None
        nat.call("qc_unpack_fock", nat.ptr(Hp), nat.ptr(y), float(scale), self.nbf, nrhs,
                 nat.ptr(F), s)
        return F[0] if squeeze else F

    # -- incremental builds (mixed path) -------------------------------------
    def set_incremental(self, tau: float | None = 1e-12, full_every: int = 0):
        """Build each Fock matrix as G(D) = G(D_prev) + G(D - D_prev), the
        difference contracted with tile screening at ``tau`` (the standard
        incremental build of direct SCF: as the SCF converges, D - D_prev
        shrinks and most tiles drop out).  A full build is made on the first
        call, after ``full_every`` incremental ones (0: never on a count),
        and from the first ``request_full_build`` on every build is full (the
        SCF drivers ask before they accept convergence, so a converged result
        always rests on full builds, and the screening errors of the
        incremental phase cannot hold the convergence test off).  Needs
        compacted supermatrices; tau=None switches it off."""
        self.inc = None if tau is None else _Incremental(float(tau), int(full_every))
        return self

    def request_full_build(self):
        if self.inc is not None:
            self.inc.force_full = True
            self.inc.final = True

    def last_build_incremental(self) -> bool:
        return self.inc is not None and self.inc.last_incremental

    def _inc_fock(self, H, dens, full_fn, inc_fn):
        """Fock matrices H + G(dens) by a full or an incremental build (see
        set_incremental); on the GPU each kind of build is one HIP graph
        replay after its first (eager) use."""
        inc = self.inc
        full = (inc.G is None or inc.force_full or inc.final
                or (inc.full_every > 0 and inc.since_full >= inc.full_every))
        kind = "full" if full else "inc"
        if inc.G is None:
            k = len(dens)
            z = torch.zeros((k,) + tuple(dens[0].shape), dtype=dens[0].dtype, device=dens[0].device)
            inc.G, inc.prev, inc.Din, inc.F = z, z.clone(), z.clone(), z.clone()
            inc.H = torch.empty_like(dens[0])
        inc.Din.copy_(torch.stack(dens))
        inc.H.copy_(H)

        def body():
            if kind == "full":
                inc.G.copy_(torch.stack(full_fn(list(inc.Din))))
            else:
                inc.G.add_(torch.stack(inc_fn(list(inc.Din - inc.prev))))
            inc.prev.copy_(inc.Din)
            inc.F.copy_(inc.H + inc.G)

        graphs = USE_FOCK_GRAPHS and inc.Din.is_cuda
        g = inc.graphs.get(kind) if graphs else None
        if g is not None:
            g.replay()
        elif graphs and kind in inc.warm:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                body()
            inc.graphs[kind] = g
            g.replay()
        else:
            body()  # first use of this kind (also the warm-up before capture)
            inc.warm.add(kind)
        if full:
            inc.since_full = 0
            inc.n_full += 1
        else:
            inc.since_full += 1
            inc.n_inc += 1
        inc.force_full = False
        inc.last_incremental = not full
        return [f.clone() for f in inc.F]

    def fock_rhf(self, H, D):
        """F = H + J(D) - 1/2 K(D) from the P supermatrix."""
        if self.inc is not None:
            P = self._compact_or_raise(self.ensure_P())
            F = self._inc_fock(H, [D], lambda d: [self.contract(P, d[0])],
                               lambda d: [self.contract(P, d[0], tau_skip=self.inc.tau)])
            return F[0]
        return self.contract(self.ensure_P(), D, H)

    def _compact_or_raise(self, M):
        if not isinstance(M, CompactSupermatrix):
            raise RuntimeError("incremental Fock builds need compacted supermatrices (use_compact)")
        return M

    def coulomb(self, D):
        """J(D); from P and K (J = P + K/2) once J has been released."""
        if self.J is None:
            if self.P is None or self.K is None:
                raise RuntimeError("J supermatrix was released")
            return self.contract(self.P, D) + self.contract(self.K, D, scale=0.5)
        return self.contract(self.J, D)

    def exchange(self, D):
        return self.contract(self.ensure_K(), D)

    def fock_uhf(self, H, Da, Db):
        """F_s = H + J(Da + Db) - K(D_s).  With P = J - K/2 assembled, as
        F_a,b = H + P(Da + Db) -+ K(Da - Db)/2 (the same matrices, K being
        linear): two tile passes with one right-hand side each instead of
        J with one and K with two, which the two-right-hand-side pass runs
        at a lower share of HBM bandwidth."""
        if self.P is not None and self.K is not None:
            return self._fock_uhf_pk(H, Da, Db)
        if self.inc is not None:
            J = self._compact_or_raise(self.J)
            K = self._compact_or_raise(self.ensure_K())

            def build(d, tau=0.0):
                Gj = self.contract(J, d[0] + d[1], tau_skip=tau)
                Gk = self.contract(K, torch.stack(d), tau_skip=tau)
                return [Gj - Gk[0], Gj - Gk[1]]

            F = self._inc_fock(H, [Da, Db], build, lambda d: build(d, self.inc.tau))
            return F[0], F[1]
        Jf = self.coulomb_plus(H, Da + Db)
        K = self.contract(self.ensure_K(), torch.stack([Da, Db]))
        return Jf - K[0], Jf - K[1]

    def _fock_uhf_pk(self, H, Da, Db):
        P, K = self.P, self.K
        if self.inc is not None:
            P, K = self._compact_or_raise(P), self._compact_or_raise(K)

        def build(d, tau=0.0, Hd=None):
            Gp = self.contract(P, d[0] + d[1], Hd, tau_skip=tau)
            Gk = self.contract(K, d[0] - d[1], scale=0.5, tau_skip=tau)
            return [Gp - Gk, Gp + Gk]

        if self.inc is not None:
            F = self._inc_fock(H, [Da, Db], build, lambda d: build(d, self.inc.tau))
            return F[0], F[1]
        return tuple(build([Da, Db], Hd=H))

    def coulomb_plus(self, H, D):
        if self.J is None:
            raise RuntimeError("J supermatrix was released")
        return self.contract(self.J, D, H)


# Replay the incremental and full Fock builds of incremental mode from HIP
# graphs after their first use (a build is ~25 launches of small kernels
# around the two tile passes); off: every build is launched eagerly.
USE_FOCK_GRAPHS = True


@dataclass
class _Incremental:
    tau: float
    full_every: int
    # device state: G(D_prev), D_prev, the input and output of a build, H
    G: torch.Tensor | None = None
    prev: torch.Tensor | None = None
    Din: torch.Tensor | None = None
    F: torch.Tensor | None = None
    H: torch.Tensor | None = None
    graphs: dict = field(default_factory=dict)
    warm: set = field(default_factory=set)
    since_full: int = 0
    force_full: bool = False
    final: bool = False
    last_incremental: bool = False
    n_full: int = 0
    n_inc: int = 0
