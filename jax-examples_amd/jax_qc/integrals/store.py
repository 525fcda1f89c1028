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
import time
from dataclasses import dataclass

import numpy as np
import torch

from jax_qc import _native as nat
from jax_qc.integrals import engine as eng


TILE_ZERO, TILE_F32, TILE_F64 = 0, 1, 2


@dataclass
class CompactSupermatrix:
    """A supermatrix re-packed tile by tile: tiles whose max |element| is
    below ``tau_zero`` are dropped, below ``tau_f32`` stored in FP32, the rest
    in FP64.  The Fock build accumulates in FP64 either way."""

    pool: torch.Tensor      # uint8 byte pool
    fmt: torch.Tensor       # uint8 per tile
    off: torch.Tensor       # int64 byte offset per tile
    n_zero: int
    n_f32: int
    n_f64: int
    bytes: int
    tau_f32: float
    tau_zero: float


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
        self.J = torch.zeros(eng.sm_numel(self.nbf), dtype=torch.float64, device=self.dev)
        eng.run_eri(self.pt, self.plan, eng.MODE_SUPER, self.J)
        torch.cuda.synchronize(self.dev)
        self.eri_seconds = time.perf_counter() - t0
        self.P = self.K = None
        self._work, self._x, self._y = {}, {}, {}

    def use_compact(self, tau_f32: float = 1e-4, tau_zero: float = 1e-14):
        """Replace the assembled supermatrices (P for RHF; J and K for UHF)
        by their compacted mixed forms.  With neither P nor K assembled yet,
        P is assembled first (the RHF case); a store that already holds K is
        taken as UHF and no P is built."""
        if self.P is None and self.K is None:
            self.ensure_P()
        if isinstance(self.P, torch.Tensor):
            self.P = self.compact(self.P, tau_f32, tau_zero)
        if isinstance(self.K, torch.Tensor):
            self.K = self.compact(self.K, tau_f32, tau_zero)
        if isinstance(self.J, torch.Tensor) and self.K is not None:
            self.J = self.compact(self.J, tau_f32, tau_zero)
        return self

    def ensure_P(self):
        if self.P is None:
            self.P = torch.empty_like(self.J)
            nat.call("qc_sm_assemble", 0, nat.ptr(self.J), nat.ptr(self.P), self.nbf,
                     nat.stream_ptr())
        return self.P

    def ensure_K(self):
        if self.K is None:
            self.K = torch.empty_like(self.J)
            nat.call("qc_sm_assemble", 1, nat.ptr(self.J), nat.ptr(self.K), self.nbf,
                     nat.stream_ptr())
        return self.K

    def release_J(self):
        """Drop the J supermatrix once only P is needed (RHF)."""
        self.ensure_P()
        self.J = None

    def compact(self, M: torch.Tensor, tau_f32: float = 1e-4, tau_zero: float = 1e-14
                ) -> CompactSupermatrix:
        """Mixed-precision, zero-skipping copy of supermatrix M."""
        T = self.T
        nt = T * (T + 1) // 2
        tmax = torch.empty(nt, dtype=torch.float64, device=self.dev)
        nat.call("qc_sm_tile_maxabs", nat.ptr(M), self.nbf, nat.ptr(tmax), nat.stream_ptr())
        tm = tmax.cpu().numpy()
        fmt = np.where(tm < tau_zero, TILE_ZERO, np.where(tm < tau_f32, TILE_F32, TILE_F64))
        fmt = fmt.astype(np.uint8)
        size = np.where(fmt == TILE_F64, 8 * 4096, np.where(fmt == TILE_F32, 4 * 4096, 0))
        off = np.zeros(nt, dtype=np.int64)
        off[1:] = np.cumsum(size)[:-1]
        total = int(size.sum())
        pool = torch.empty(max(16, total), dtype=torch.uint8, device=self.dev)
        fd = torch.from_numpy(fmt).to(self.dev)
        od = torch.from_numpy(off).to(self.dev)
        nat.call("qc_sm_pack", nat.ptr(M), self.nbf, nat.ptr(fd), nat.ptr(od), nat.ptr(pool),
                 nat.stream_ptr())
        return CompactSupermatrix(pool, fd, od, int((fmt == 0).sum()), int((fmt == 1).sum()),
                                  int((fmt == 2).sum()), total, tau_f32, tau_zero)

    # -- Fock builds --------------------------------------------------------
    def bytes_per_matrix(self) -> int:
        return eng.sm_numel(self.nbf) * 8

    def _buffers(self, nrhs):
        if nrhs not in self._work:
            wb = nat.lib().qc_symv_workspace_bytes(self.nbf, nrhs)
            self._work[nrhs] = torch.empty(max(1, wb // 8), dtype=torch.float64, device=self.dev)
            self._x[nrhs] = torch.empty(nrhs * self.npad, dtype=torch.float64, device=self.dev)
            self._y[nrhs] = torch.empty(nrhs * self.npad, dtype=torch.float64, device=self.dev)
        return self._work[nrhs], self._x[nrhs], self._y[nrhs]

    def contract(self, M: torch.Tensor, D: torch.Tensor, H: torch.Tensor | None = None,
                 scale: float = 1.0) -> torch.Tensor:
        """Return H + scale * M D~ for a density or a stack of densities."""
        D = D.contiguous()
        squeeze = D.dim() == 2
        if squeeze:
            D = D.unsqueeze(0)
        nrhs = D.shape[0]
        work, x, y = self._buffers(nrhs)
        s = nat.stream_ptr()
        nat.call("qc_pack_density", nat.ptr(D), self.nbf, nrhs, nat.ptr(x), s)
        if isinstance(M, CompactSupermatrix):
            nat.call("qc_symv_mixed_tiles", nat.ptr(M.pool), nat.ptr(M.fmt), nat.ptr(M.off),
                     self.nbf, nrhs, nat.ptr(x), nat.ptr(work), s)
            nat.call("qc_symv_reduce", self.nbf, nrhs, nat.ptr(y), nat.ptr(work), s)
        else:
            nat.call("qc_symv", nat.ptr(M), self.nbf, nrhs, nat.ptr(x), nat.ptr(y), nat.ptr(work), s)
        F = torch.empty_like(D)
        Hp = H.contiguous() if H is not None else None
        nat.call("qc_unpack_fock", nat.ptr(Hp), nat.ptr(y), float(scale), self.nbf, nrhs,
                 nat.ptr(F), s)
        return F[0] if squeeze else F

    def fock_rhf(self, H, D):
        """F = H + J(D) - 1/2 K(D) from the P supermatrix."""
        return self.contract(self.ensure_P(), D, H)

    def coulomb(self, D):
        if self.J is None:
            raise RuntimeError("J supermatrix was released")
        return self.contract(self.J, D)

    def exchange(self, D):
        return self.contract(self.ensure_K(), D)

    def fock_uhf(self, H, Da, Db):
        """F_s = H + J(Da + Db) - K(D_s)."""
        Jf = self.coulomb_plus(H, Da + Db)
        K = self.contract(self.ensure_K(), torch.stack([Da, Db]))
        return Jf - K[0], Jf - K[1]

    def coulomb_plus(self, H, D):
        if self.J is None:
            raise RuntimeError("J supermatrix was released")
        return self.contract(self.J, D, H)
