"""Pins the oracle's Boys functions (both modes) against mpmath and the
reference's own Boys tests (tests/test_boys.py:22-75)."""

import json
import math
import os

import numpy as np
import pytest

from oracle import mmd

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "boys_mpmath.json")))


def test_accurate_boys_matches_mpmath():
    t = np.array(GOLD["t"])
    F = np.array(GOLD["F"])
    for n in range(GOLD["nmax"] + 1):
        got = mmd.boys(n, t, mmd.BOYS_ACCURATE)
        np.testing.assert_allclose(got, F[:, n], rtol=2e-15, atol=1e-300)


def test_reference_boys_f0_closed_form():
    # test_boys.py:22-34: F_0(0) = 1; series branch and erf branch agree.
    assert mmd.boys(0, 0.0, mmd.BOYS_REFERENCE)[0] == pytest.approx(1.0, abs=1e-14)
    for t in (1e-6, 1e-4, 1e-3, 5e-3):
        ref = 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))
        assert mmd.boys(0, t, mmd.BOYS_REFERENCE)[0] == pytest.approx(ref, abs=1e-14)
    for t in (0.5, 1.0, 5.0, 20.0, 100.0):
        ref = 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))
        assert mmd.boys(0, t, mmd.BOYS_REFERENCE)[0] == pytest.approx(ref, rel=1e-12, abs=1e-14)


def test_reference_boys_fn_at_zero_and_recurrence():
    # test_boys.py:58-70
    for n in range(5):
        assert mmd.boys(n, 0.0, mmd.BOYS_REFERENCE)[0] == pytest.approx(1 / (2 * n + 1), abs=1e-14)
    for t in (0.1, 1.0, 5.0):
        f0 = 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))
        f1 = (f0 - math.exp(-t)) / (2 * t)
        assert mmd.boys(1, t, mmd.BOYS_REFERENCE)[0] == pytest.approx(f1, rel=1e-12, abs=1e-14)


def test_reference_boys_is_accurate_where_stable():
    # The reference's upward recursion is accurate for low n, and for any n
    # once t is large; it is not near t = 0.01 for n >= 4 (documented defect).
    t = np.array(GOLD["t"])
    F = np.array(GOLD["F"])
    for n, rtol in ((0, 1e-12), (1, 1e-11), (2, 1e-10), (3, 5e-9)):
        got = mmd.boys(n, t, mmd.BOYS_REFERENCE)
        np.testing.assert_allclose(got, F[:, n], rtol=rtol, atol=1e-14)
    big = t >= 20.0
    for n in range(9):
        np.testing.assert_allclose(mmd.boys(n, t[big], mmd.BOYS_REFERENCE), F[big, n], rtol=1e-10)


def test_reference_boys_defect_is_real():
    # the cancellation window the oracle's accurate mode exists for
    ref = mmd.boys(8, 0.0101, mmd.BOYS_REFERENCE)[0]
    acc = mmd.boys(8, 0.0101, mmd.BOYS_ACCURATE)[0]
    assert abs(ref - acc) > 1.0
