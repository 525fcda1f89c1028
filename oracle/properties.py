"""Test infrastructure (not the product): one-electron moment integrals by
Gauss-Hermite quadrature, independent of the Hermite E coefficients used by
the device kernels.

For a primitive pair the 3-D integral factorises into 1-D integrals
  I(a, b, k) = int (x - A)^a (x - B)^b (x - O)^k exp(-al (x-A)^2 - be (x-B)^2) dx
             = exp(-al be/p (A-B)^2) int poly(x) exp(-p (x - P)^2) dx,
evaluated exactly by an 8-point Gauss-Hermite rule (exact to polynomial
degree 15).  Restricted to s and p shells, whose spherical and Cartesian
orders coincide (px, py, pz; jax_qc/integrals/mcmurchie_davidson.py:39-50),
so no spherical transform is involved.  The contraction coefficients are the
normalised ones of the basis (jax_qc/basis/normalize.py:36-107), as in the
reference's overlap (jax_qc/integrals/overlap.py:67-85).

The reference's own dipole integrals (jax_qc/properties/dipole.py:26-89) are
an approximation (overlap times an exponent-weighted centre), so they are not
what this checks against; this is the exact moment integral.
"""

from __future__ import annotations

import numpy as np

_GH_X, _GH_W = np.polynomial.hermite.hermgauss(8)


def _cart(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def _moment_1d(a, A, ia, b, B, ib, O, k):
    p = a + b
    P = (a * A + b * B) / p
    x = P + _GH_X / np.sqrt(p)
    f = (x - A) ** ia * (x - B) ** ib * (x - O) ** k
    return np.exp(-a * b / p * (A - B) ** 2) * np.sum(_GH_W * f) / np.sqrt(p)


def moment_matrices(basis, origin=(0.0, 0.0, 0.0)):
    """(S, Mx, My, Mz): overlap and first moments about ``origin`` for a
    basis of s and p shells, AO order of the basis."""
    sh = basis.shells
    if any(s.angular_momentum > 1 for s in sh):
        raise NotImplementedError("quadrature restatement covers s and p shells only")
    O = np.asarray(origin, dtype=np.float64)
    off = np.concatenate([[0], np.cumsum([2 * s.angular_momentum + 1 for s in sh])])
    n = int(off[-1])
    out = np.zeros((4, n, n))
    for i, si in enumerate(sh):
        for j, sj in enumerate(sh):
            A, B = np.asarray(si.center, float), np.asarray(sj.center, float)
            for ca, (ax, ay, az) in enumerate(_cart(si.angular_momentum)):
                for cb, (bx, by, bz) in enumerate(_cart(sj.angular_momentum)):
                    acc = np.zeros(4)
                    for a, c1 in zip(si.exponents, si.coefficients):
                        for b, c2 in zip(sj.exponents, sj.coefficients):
                            s = [_moment_1d(a, A[d], (ax, ay, az)[d], b, B[d], (bx, by, bz)[d],
                                            O[d], 0) for d in range(3)]
                            m = [_moment_1d(a, A[d], (ax, ay, az)[d], b, B[d], (bx, by, bz)[d],
                                            O[d], 1) for d in range(3)]
                            w = c1 * c2
                            acc[0] += w * s[0] * s[1] * s[2]
                            acc[1] += w * m[0] * s[1] * s[2]
                            acc[2] += w * s[0] * m[1] * s[2]
                            acc[3] += w * s[0] * s[1] * m[2]
                    out[:, off[i] + ca, off[j] + cb] = acc
    return out
