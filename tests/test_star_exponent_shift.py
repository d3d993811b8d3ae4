"""The mixed star pass's stream exponents, pinned on the CPU.

Per star and stream the convolution sums kw3_j 2^-(t_j^2 + e) over the
window's points, t_j = (p r_star) ratio_j - m.  The kernel takes each
exponent in FP64, shifted by t0^2, the smallest t^2 over the window (at an
end, or 0 where the Gaussian's peak lies inside), rounds the shifted value to
FP32 for the hardware 2^x and multiplies the sum by 2^-(e + t0^2) in FP64.
Where the peak lies inside the window and p r_star exceeds 2 / (the widest
gap between adjacent ratio_j), the points may straddle the peak with every
exponent above 1, and the kernel splits each exponent into an integer and a
fraction in FP64 instead, only the fraction going through FP32.
These tests model that arithmetic (the FP32 rounding of what enters the
2^x, with 2^x itself exact) over sights whose Gaussians sit anywhere around
the window, narrow or wide against the point spacing, and hold the sum's
relative error inside the mixed pass's per-star budget of 1e-7.
Either half of the rule alone fails it.
"""
import math

import numpy as np
import pytest


def conv_tables(n):
    """ratio_j = 10^(dx_j / 5) and kw3_j = kw_j ratio_j^3 over the reference's
    Gauss-Legendre offsets (+-3 x 0.6 mag)."""
    x, wg = np.polynomial.legendre.leggauss(n)
    dx = 3 * 0.6 * x
    kw = wg * np.exp(-dx * dx / (2 * 0.36)) / (0.6 * math.sqrt(2 * math.pi))
    ratio = 10 ** (dx / 5)
    return ratio, kw * ratio ** 3


def shifted_sum(ratio, kw3, pr, m, window_shift=True, split=True):
    t = pr * ratio - m
    t0 = 0.0
    if window_shift:
        lo, hi = min(t[0], t[-1]), max(t[0], t[-1])
        t0 = lo if lo > 0 else (hi if hi < 0 else 0.0)
    # the kernel forms t^2 - t0^2 by one FP64 FMA; the difference from
    # this FP64 expression is far below the FP32 rounding modelled
    a = t * t - t0 * t0
    coarse_pr = 2.0 / np.max(np.abs(np.diff(ratio)))
    if split and t0 == 0.0 and pr > coarse_pr:
        k = np.floor(a)
        terms = np.exp2(-np.float32(a - k).astype(np.float64) - k)
    else:
        terms = np.exp2(-np.float32(a).astype(np.float64))
    return np.sum(kw3 * terms) * 2.0 ** (-t0 * t0)


def worst_error(cases=6000, sizes=(30, 60, 120), **rule):
    rng = np.random.default_rng(20261203)
    worst = 0.0
    for n in sizes:
        ratio, kw3 = conv_tables(n)
        for _ in range(cases):
            pr = math.exp(rng.uniform(math.log(0.05), math.log(1000.0)))
            m = pr * rng.uniform(0.3, 3.0) + rng.uniform(-10.0, 10.0)
            t = pr * ratio - m
            exact = np.sum(kw3 * np.exp2(-t * t))
            if not exact > 1e-280:
                continue
            got = shifted_sum(ratio, kw3, pr, m, **rule)
            worst = max(worst, abs(got / exact - 1.0))
    return worst


@pytest.mark.parametrize("n, bound", [(30, 1e-7), (60, 5e-8), (120, 5e-8)])
def test_exponent_rule_bound(n, bound):
    """Relative error of a star's stream sum from the exponents' rounding:
    7.9e-8 measured at 30 points, 4.5e-8 at 60 and 3.6e-8 at 120 (the reference's
    parameter files convolve over 120).  The worst sights have the peak just
    outside the window, where the weights' taper puts the terms that carry
    the sum a few units of exponent inside it; the FP32 2^x adds its own
    1-ulp error to every term in either case."""
    worst = worst_error(sizes=(n,))
    print(f"exponent rule, {n} points: worst {worst:.2e}")
    assert worst <= bound


def test_sight_minimum_alone_fails():
    """Without the window's shift the rounding reaches 1e-6 and beyond."""
    worst = worst_error(cases=2000, window_shift=False)
    print(f"sight minimum alone: worst {worst:.2e}")
    assert worst > 1e-6


def test_window_shift_alone_fails():
    """Without the split for peaks between distant points: above 1e-6."""
    worst = worst_error(cases=2000, split=False)
    print(f"window shift alone: worst {worst:.2e}")
    assert worst > 1e-6
