"""Analytic nuclear gradients from the device kernels.

dE/dR = dE_nuc/dR + sum D d(T+V)/dR - sum W dS/dR + 1/2 sum G d(ERI)/dR,
the Hellmann-Feynman + Pulay expression the reference evaluates in
jax_qc/grad/rhf_grad.py:186-209 -- there with every derivative a central
finite difference of full integral rebuilds at fixed D and W; here each
term is contracted analytically on the GPU (qc_grad1e, qc_grad2e) without
forming derivative integrals.  Valid for converged RHF and UHF states
(for UHF: D = Da + Db, W = Wa + Wb, and the exchange term uses each spin
density -- the reference applies its RHF expression to UHF results, which
is not the UHF gradient).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from jax_qc import _native as nat
from jax_qc.integrals import engine as eng


# Kernel form per (bra class, ket class) of the derivative-ERI contraction:
# 1 one thread per quartet, 2 one wave per quartet (bra side contracted with
# G first), 3 one lane per primitive quartet, 4 form 2 with the G contraction
# on the matrix cores, 5 one lane per primitive quartet with G contracted with
# the bra E first (the default where it exists), 0 the library's default.
# Classes absent from the table use the default; a probe may set
# GRAD_VARIANT_OVERRIDE to force one form for every class that has it.
GRAD_VARIANTS: dict = {}
GRAD_VARIANT_OVERRIDE = None
# pass the atom count to qc_grad2e so that the thread-per-quartet kernels sum
# a work-group's contributions in LDS (False: every add goes to the global
# gradient, the form used above 256 atoms)
GRAD_STAGE_IN_LDS = True


def grad_variant_for(cab: int, ccd: int, fp32: int = 0) -> int:
    if GRAD_VARIANT_OVERRIDE is not None:
        ov = GRAD_VARIANT_OVERRIDE
        if ov == 1 and (cab, ccd) not in THREAD_CAPABLE:
            return 2
        if ov == 3 and (cab, ccd) not in (PRIM_CAPABLE_F32 if fp32 else PRIM_CAPABLE):
            return 2
        if ov == 4 and (cab, ccd) not in WAVE_MFMA_CAPABLE:
            return 2
        if ov == 5 and (cab, ccd) not in (YPRIM_CAPABLE_F32 if fp32 else YPRIM_CAPABLE):
            return 2
        return ov
    return GRAD_VARIANTS.get((cab, ccd, int(fp32)), GRAD_VARIANTS.get((cab, ccd), 0))


# classes with a one-thread-per-quartet kernel (as built by csrc/build.py)
THREAD_CAPABLE = {(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (3, 0), (3, 1), (4, 0), (3, 3)}
# classes with the one-lane-per-primitive-quartet kernel (csrc/build.py
# GRAD_PRIM, GRAD_PRIM_F32): (pp|pp), (ds|pp); in FP32 also (dp|ps), (dp|pp),
# (dp|ds)
PRIM_CAPABLE = {(2, 2), (3, 2)}
PRIM_CAPABLE_F32 = PRIM_CAPABLE | {(4, 1), (4, 2), (4, 3)}
# classes with the one-wave-per-quartet kernel whose Y phase (G contracted
# with the bra E products) runs on the matrix cores (variant 4; csrc/build.py
# GRAD_WAVE_MFMA): (pp|pp), (ds|pp), (dp|ps), (dp|pp), (dp|dp)
WAVE_MFMA_CAPABLE = {(2, 2), (3, 2), (4, 1), (4, 2), (4, 4)}
# classes with the one-lane-per-primitive-quartet kernel that contracts G with
# the bra E first (variant 5; csrc/build.py GRAD_YPRIM)
YPRIM_CAPABLE = {(2, 2), (3, 2), (4, 1), (4, 2), (4, 3), (4, 4), (5, 0), (5, 1), (5, 2), (5, 3),
                 (5, 4)}
YPRIM_CAPABLE_F32 = YPRIM_CAPABLE - {(5, 0)}


def energy_weighted_density(C, orbital_energies, n_occ: int, factor: float = 2.0):
    """W = factor * sum_{i < n_occ} eps_i C_i C_i^T (rhf_grad.py:92-103)."""
    Co = C[:, :n_occ]
    return factor * (Co * orbital_energies[:n_occ][None, :]) @ Co.T


def nuclear_repulsion_gradient(mol) -> np.ndarray:
    c = np.asarray(mol.coords, dtype=np.float64)
    Z = np.asarray(mol.atomic_numbers, dtype=np.float64)
    d = c[:, None, :] - c[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    np.fill_diagonal(r, np.inf)
    return -np.sum((Z[:, None] * Z[None, :] / r ** 3)[:, :, None] * d, axis=1)


def _densities(mol, res):
    """(Dj, [Dk...], cx, W) in the spherical AO basis."""
    if res.uhf is not None:
        u = res.uhf
        Da, Db = u.density_alpha, u.density_beta
        Wa = energy_weighted_density(u.coefficients_alpha, u.orbital_energies_alpha, mol.n_alpha, 1.0)
        Wb = energy_weighted_density(u.coefficients_beta, u.orbital_energies_beta, mol.n_beta, 1.0)
        return Da + Db, [Da, Db], 0.5, Wa + Wb
    D = res.state.density
    W = energy_weighted_density(res.state.coefficients, res.state.orbital_energies,
                                mol.n_electrons // 2, 2.0)
    return D, [D], 0.25, W


# Drop the quartets that the kernels' density screen would skip before the
# launch (see prescreen_launches); the gradient is the same either way.
PRESCREEN = True


def _shell_block_max(M, pt, shell_of_cart):
    """(nshell, nshell) max |M| over each pair of shells' Cartesian block."""
    ns = pt.nshell
    idx = (shell_of_cart[:, None] * ns + shell_of_cart[None, :]).reshape(-1)
    out = torch.zeros(ns * ns, dtype=M.dtype, device=M.device)
    out.scatter_reduce_(0, idx, M.abs().reshape(-1), reduce="amax")
    return out.reshape(ns, ns)


def prescreen_launches(launches, pt, shells, Mj, Mk, cx, Qd, tau_density):
    """Copies of the class launches without the quartets (ab|cd) whose bound
    (Mj[a,b] Mj[c,d] + cx sum_s (Mk_s[a,c] Mk_s[b,d] + Mk_s[a,d] Mk_s[b,c]))
    * Q_ab Q_cd, with Mj, Mk_s the shell-block maxima of |Dj|, |Dk_s|, is
    below tau_density / (1 + 1e-6): that bound is at least the max |G_abcd|
    the kernels compute (in FP64 or FP32), so each quartet left out is one
    the kernel's own density screen skips, and the gradient is unchanged --
    only the waves that would have exited after forming G are not launched.
    qc_grad2e_prescreen counts every launch's kept kets per bra (one wave per
    bra pair); the counts are read back together (one host synchronisation
    for the whole gradient), and a second pass writes the kept kets of each
    bra, in list order, at its prefix-sum offset.  Launches left empty are
    dropped."""
    from jax_qc.integrals.engine import ClassLaunch

    staged = [L for L in launches if L.nquartets > 0]
    if not staged:
        return []
    if not 1 <= len(Mk) <= 2:
        raise ValueError(f"prescreen_launches: {len(Mk)} exchange densities (1 or 2 expected)")
    dev = staged[0].bra.device
    sh32 = shells.to(torch.int32).contiguous()
    Mj = Mj.contiguous()
    Mk = [m.contiguous() for m in Mk]
    Mk1 = Mk[1] if len(Mk) > 1 else None
    Qd = Qd.contiguous()
    ns = int(Mj.shape[0])
    s = nat.stream_ptr()

    def run(L, count, write_off, out_ket):
        nat.call("qc_grad2e_prescreen", L.nbra, nat.ptr(L.offsets), nat.ptr(L.bra), nat.ptr(L.ket),
                 nat.ptr(L.kstart), nat.ptr(sh32), ns, nat.ptr(Mj), nat.ptr(Mk[0]), nat.ptr(Mk1),
                 float(cx), nat.ptr(Qd), float(tau_density), nat.ptr(count), nat.ptr(write_off),
                 nat.ptr(out_ket), s)

    counts = []
    for L in staged:
        cnt = torch.empty(L.nbra, dtype=torch.int64, device=dev)
        run(L, cnt, None, None)
        counts.append(cnt)
    kept = torch.stack([c.sum() for c in counts]).cpu().tolist()
    out = []
    for L, cnt, n in zip(staged, counts, kept):
        n = int(n)
        if n == 0:
            continue
        offs = torch.zeros(L.nbra + 1, dtype=torch.int64, device=dev)
        offs[1:] = torch.cumsum(cnt, 0)
        ket = torch.empty(n, dtype=torch.int32, device=dev)
        run(L, None, offs, ket)
        out.append(ClassLaunch(L.cab, L.ccd, L.fp32, L.nbra, n, L.bra, ket, offs,
                               offs[:-1].to(torch.int32).contiguous()))
    return out


def electronic_gradient(mol, basis, Dj, Dk, cx, W, tau: float = 1e-13, pair_table=None,
                        parts=("1e", "2e"), tau_density: float = 0.0,
                        tau32: float | None = None, report: list | None = None) -> torch.Tensor:
    """sum D d(T+V) - sum W dS + 1/2 sum G dERI on the device, (natom, 3).
    tau: Schwarz threshold of the quartet list; tau_density > 0 also skips
    quartets with max|G_abcd| Q_ab Q_cd below it (density-weighted); tau32:
    quartets with Schwarz bound below it sum their primitives in FP32 (the
    contributions are added in FP64)."""
    dev = nat.device()
    nat.ensure_init(dev)
    pt = pair_table or eng.PairTable(basis, dev)
    Cm = torch.from_numpy(eng.cart_to_sph_matrix(basis)).to(dev)

    def cart(M):
        return (Cm.T @ M @ Cm).contiguous()

    Djc = cart(Dj)
    Dkc = [cart(x) for x in Dk]
    Wc = cart(W)
    nat_ = len(mol.atomic_numbers)
    grad = torch.zeros(nat_ * 3, dtype=torch.float64, device=dev)
    s = nat.stream_ptr()
    if "1e" in parts:
        pt1 = getattr(pt, "unscreened", pt)  # every primitive for the 1e terms
        xyz = torch.as_tensor(np.ascontiguousarray(mol.coords, dtype=np.float64), device=dev)
        zz = torch.as_tensor(np.asarray(mol.atomic_numbers, dtype=np.float64), device=dev)
        for c in range(6):
            lst = pt1.class_pairs[c]
            if len(lst) == 0:
                continue
            ld = torch.from_numpy(lst).to(dev)
            nat.call("qc_grad1e", c, nat.ptr(pt1.pairs64), nat.ptr(pt1.sp_prim_off),
                     nat.ptr(pt1.sp_nprim), nat.ptr(pt1.sp_aoc), nat.ptr(pt1.sp_atom),
                     nat.ptr(pt1.sp_diag), nat.ptr(ld), len(lst), nat_, nat.ptr(xyz), nat.ptr(zz),
                     pt1.ncart, nat.ptr(Djc), nat.ptr(Wc), nat.ptr(grad), s)
    if "2e" in parts:
        plan = eng.plan_quartets(pt, tau, tau32)
        Qd = torch.from_numpy(pt.Q).to(dev) if tau_density > 0.0 else None
        pre = None
        if Qd is not None and PRESCREEN:
            soc = torch.from_numpy(np.repeat(np.arange(pt.nshell),
                                             np.diff(pt.cart_off)).astype(np.int64)).to(dev)
            pre = (torch.from_numpy(pt.sp_shells.astype(np.int64)).to(dev),
                   _shell_block_max(Djc, pt, soc), [_shell_block_max(x, pt, soc) for x in Dkc])
        launches = plan.launches
        if pre is not None:
            launches = prescreen_launches(launches, pt, pre[0], pre[1], pre[2], cx, Qd, tau_density)
        for L in launches:
            a = nat.QcGrad2eArgs()
            a.pairs64 = pt.pairs64.data_ptr()
            a.sp_prim_off = pt.sp_prim_off.data_ptr()
            a.sp_nprim = pt.sp_nprim.data_ptr()
            a.sp_aoc = pt.sp_aoc.data_ptr()
            a.sp_atom = pt.sp_atom.data_ptr()
            a.sp_diag = pt.sp_diag.data_ptr()
            a.bra_list = L.bra.data_ptr()
            a.ket_list = L.ket.data_ptr()
            a.offsets = L.offsets.data_ptr()
            a.kstart = L.kstart.data_ptr()
            a.nbra = L.nbra
            a.nquartets = L.nquartets
            a.ncart = pt.ncart
            a.dj = Djc.data_ptr()
            a.dk0 = Dkc[0].data_ptr()
            a.dk1 = Dkc[1].data_ptr() if len(Dkc) > 1 else 0
            a.nk = len(Dkc)
            a.cx = cx
            a.grad = grad.data_ptr()
            a.natom = nat_ if GRAD_STAGE_IN_LDS else 0
            a.variant = grad_variant_for(L.cab, L.ccd, L.fp32)
            a.fp32 = L.fp32
            if L.fp32:
                a.pairs32 = pt.pairs32.data_ptr()
            if Qd is not None:
                a.schwarz = Qd.data_ptr()
                a.tau_density = tau_density
            if report is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            nat.call("qc_grad2e", L.cab, L.ccd, ctypes.byref(a), s)
            if report is not None:
                e1.record()
                report.append(dict(cab=L.cab, ccd=L.ccd, fp32=L.fp32, nquartets=L.nquartets,
                                   variant=a.variant, events=(e0, e1), bra=L.bra, ket=L.ket,
                                   offsets=L.offsets, kstart=L.kstart))
    return grad.reshape(nat_, 3)


def scf_gradient(mol, basis, res, tau: float = 1e-13, pair_table=None,
                 tau_density: float = 0.0, tau32: float | None = None) -> torch.Tensor:
    """Full nuclear gradient (Eh/Bohr) of a converged RHF or UHF result."""
    Dj, Dk, cx, W = _densities(mol, res)
    g = electronic_gradient(mol, basis, Dj, Dk, cx, W, tau, pair_table,
                            tau_density=tau_density, tau32=tau32)
    return g + torch.from_numpy(nuclear_repulsion_gradient(mol)).to(g.device)
