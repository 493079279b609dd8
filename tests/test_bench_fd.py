"""The finite-difference check of the adjoint (bench.py's fd_check) on
synthetic piecewise-linear functions with rounding: its bar covers the
difference between the central difference and the slope of the piece that
holds s = 0 when one kink lies within the step, and shrinks to the rounding
of J on a straight line."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _kinked(x, kinks, slopes):
    """J(x) - J(0) for slope slopes[i] between kinks[i-1] and kinks[i]
    (kinks sorted), by exact integration of the slope from 0 to x."""
    pts = sorted([0.0, x] + [k for k in kinks if min(0.0, x) < k < max(0.0, x)])
    v = 0.0
    for a, b in zip(pts[:-1], pts[1:]):
        v += slopes[int(np.searchsorted(kinks, 0.5 * (a + b)))] * (b - a)
    return v if x >= 0 else -v


def test_one_kink_within_the_step_is_inside_the_bar():
    h = 5e-10
    base = 1.24e6
    for s_kink, right in ((0.3 * h, 1.5e3), (-0.7 * h, -4e2), (0.9 * h, 1e3)):
        kinks = [s_kink]
        slopes = [1.1e3, 1.1e3 + right] if s_kink > 0 else [1.1e3 + right, 1.1e3]
        J = [base + _kinked(x, kinks, slopes) for x in (-h, 0.0, h)]
        c, bar = bench.fd_check(*J, h)
        assert abs(c - 1.1e3) > 1.0                 # the difference is off ...
        assert abs(c - 1.1e3) <= bar                # ... and its bar says so


def test_a_straight_line_has_the_rounding_bar():
    h = 5e-10
    J = [1.24e6 + 1108.5 * x for x in (-h, 0.0, h)]
    c, bar = bench.fd_check(*J, h)
    ulp = float(np.spacing(1.24e6))
    assert bar == 2 * ulp / h
    assert abs(c - 1108.5) <= bar
