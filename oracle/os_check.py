"""Independent primitive ERI by Obara-Saika vertical recursion + horizontal
recursion (test infrastructure only).

This is a different algorithm from McMurchie-Davidson (no Hermite
expansion): it exists to pin the MMD restatement in ``oracle.mmd``.
Reference: Obara & Saika, J. Chem. Phys. 84, 3963 (1986); Head-Gordon &
Pople, J. Chem. Phys. 89, 5777 (1988).  Boys values come from the accurate
series (mpmath-checked), so a shared Boys mistake cannot hide.
"""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from oracle import mmd


def _cart(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def primitive_eri(a, A, la, b, B, lb, c, C, lc, d, D, ld):
    A, B, C, D = (np.asarray(x, dtype=np.float64) for x in (A, B, C, D))
    p, q = a + b, c + d
    P, Q = (a * A + b * B) / p, (c * C + d * D) / q
    W = (p * P + q * Q) / (p + q)
    rho = p * q / (p + q)
    Kab = math.exp(-a * b / p * float(np.dot(A - B, A - B)))
    Kcd = math.exp(-c * d / q * float(np.dot(C - D, C - D)))
    T = rho * float(np.dot(P - Q, P - Q))
    L = la + lb + lc + ld
    F = [float(mmd.boys(m, T)[0]) for m in range(L + 1)]
    base = 2.0 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd
    PA, WP, QC, WQ = P - A, W - P, Q - C, W - Q

    @lru_cache(maxsize=None)
    def vrr(e, f, m):
        # [e 0 | f 0]^(m), e and f Cartesian exponent tuples
        if any(x < 0 for x in e) or any(x < 0 for x in f):
            return 0.0
        if sum(e) == 0 and sum(f) == 0:
            return base * F[m]
        if sum(e) > 0:
            i = next(k for k in range(3) if e[k] > 0)
            em = tuple(e[k] - (k == i) for k in range(3))
            emm = tuple(e[k] - 2 * (k == i) for k in range(3))
            fm = tuple(f[k] - (k == i) for k in range(3))
            v = PA[i] * vrr(em, f, m) + WP[i] * vrr(em, f, m + 1)
            if em[i] > 0:
                v += em[i] / (2 * p) * (vrr(emm, f, m) - rho / p * vrr(emm, f, m + 1))
            if f[i] > 0:
                v += f[i] / (2 * (p + q)) * vrr(em, fm, m + 1)
            return v
        i = next(k for k in range(3) if f[k] > 0)
        fm = tuple(f[k] - (k == i) for k in range(3))
        fmm = tuple(f[k] - 2 * (k == i) for k in range(3))
        v = QC[i] * vrr(e, fm, m) + WQ[i] * vrr(e, fm, m + 1)
        if fm[i] > 0:
            v += fm[i] / (2 * q) * (vrr(e, fmm, m) - rho / q * vrr(e, fmm, m + 1))
        return v

    AB, CD = A - B, C - D

    @lru_cache(maxsize=None)
    def hrr(ea, eb, ec, ed):
        # (ea eb | ec ed) by transferring angular momentum from b to a, d to c
        if sum(eb) > 0:
            i = next(k for k in range(3) if eb[k] > 0)
            ebm = tuple(eb[k] - (k == i) for k in range(3))
            eap = tuple(ea[k] + (k == i) for k in range(3))
            return hrr(eap, ebm, ec, ed) + AB[i] * hrr(ea, ebm, ec, ed)
        if sum(ed) > 0:
            i = next(k for k in range(3) if ed[k] > 0)
            edm = tuple(ed[k] - (k == i) for k in range(3))
            ecp = tuple(ec[k] + (k == i) for k in range(3))
            return hrr(ea, eb, ecp, edm) + CD[i] * hrr(ea, eb, ec, edm)
        return vrr(ea, ec, 0)

    ca, cb, cc, cd = _cart(la), _cart(lb), _cart(lc), _cart(ld)
    out = np.zeros((len(ca), len(cb), len(cc), len(cd)))
    for i, x in enumerate(ca):
        for j, y in enumerate(cb):
            for k, z in enumerate(cc):
                for l, w in enumerate(cd):
                    out[i, j, k, l] = hrr(x, y, z, w)
    return out
