"""Gaussian normalisation (host side, construction time only).

Same convention as jax_qc/basis/normalize.py:36-107: every Cartesian
component of a shell shares the (l,0,0) primitive norm
N = (2a/pi)^(3/4) sqrt((4a)^l / (2l-1)!!), and the contracted shell is then
scaled to unit self-overlap.  The spherical d rows carry the sqrt(3) factors
that compensate for the shared norm (spherical.py:44-62).
"""

from __future__ import annotations

import math

import numpy as np


def double_factorial(n: int) -> int:
    out = 1
    while n > 1:
        out *= n
        n -= 2
    return out


def primitive_norm(alpha: float, l: int) -> float:
    base = (2.0 * alpha / math.pi) ** 0.75
    if l == 0:
        return float(base)
    return float(base * math.sqrt((4.0 * alpha) ** l / double_factorial(2 * l - 1)))


def contraction_norm(exponents, coeffs, l: int) -> float:
    a = np.asarray(exponents, dtype=np.float64)
    c = np.asarray(coeffs, dtype=np.float64)
    s = a[:, None] + a[None, :]
    ang = double_factorial(2 * l - 1) / (2.0 * s) ** l if l > 0 else 1.0
    total = float(np.sum(c[:, None] * c[None, :] * (math.pi / s) ** 1.5 * ang))
    if total <= 0.0:
        raise ValueError(f"Non-positive self-overlap ({total}) for contracted shell with l={l}.")
    return 1.0 / math.sqrt(total)


def normalize_shell(exponents, coeffs, l: int) -> np.ndarray:
    a = np.asarray(exponents, dtype=np.float64)
    d = np.asarray(coeffs, dtype=np.float64)
    c = d * np.array([primitive_norm(float(x), l) for x in a])
    return c * contraction_norm(a, c, l)
