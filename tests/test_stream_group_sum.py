"""The mixed integral pass's stream sums on the exact tables, pinned on the CPU.

Per sight and stream the pass sums qw_ij 2^-(t_ij^2 - s2) over every r-step
i and convolution point j of the region, t_ij = p r_ij - m and s2 the
smallest t^2 over the region's distances (evaluation_optimized.c:222-296 per
point, as the sight's shifted quadratic).  The kernel forms each exponent in
FP64 and rounds it to FP32 for the hardware 2^x.  In rows of at least 100
convolution points it sums groups of 8 consecutive points of one r-step in
FP32 with the weight qw rounded to FP32 (one FMA per term) and adds each
group's sum in FP64; coarser rows add every term in FP64.
These tests model that arithmetic (2^x exact, then rounded to FP32; the FMA
as the exact product and sum rounded to FP32) and compare it with the same
sums in FP64:

* per sight, over streams narrow and wide, inside and beyond the region;
* in an integral: sums over bundles of sights crossing one stream, where the
  FP32 weights' rounding, the same in every column, averages only over the
  points of a row.  At 120 points per row the groups leave a few 1e-9; at
  8, 30 or 64 points about 1e-8, which is why coarse rows stay on FP64
  terms (near 1e-9 and below there).
* An FP32 sum drops terms under half an ulp of its running sum, always
  downwards.  Short groups keep that bias small; summing each r-step's whole
  row in FP32 instead drifts, and fails the bias bound.
"""
import math

import numpy as np
import pytest

GROUP = 8
GROUP_MIN_CONV = 100


def tables(R, n):
    """r_ij and qw_ij for R r-steps uniform in magnitude over [16, 22.5] and
    n Gauss-Legendre convolution points over +-3 x 0.6 mag (the reference's
    r_kpc = 10^((g - 14.2) / 5) and qw_r3_N = W_j r_ij^3 N(dx_j))."""
    x, wg = np.polynomial.legendre.leggauss(n)
    dx = 3 * 0.6 * x
    kw = wg * np.exp(-dx * dx / (2 * 0.36)) / (0.6 * math.sqrt(2 * math.pi))
    g = np.linspace(16.0, 22.5, R + 1)
    rr = 10 ** ((0.5 * (g[1:] + g[:-1]) - 14.2) / 5)
    r = rr[:, None] * 10 ** (dx / 5)[None, :]
    return r, kw[None, :] * r ** 3


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def column_sums(r, qw, p, m, group):
    """FP64 sums and the kernel's for sights (p, m), scaled back by 2^-s2:
    group > 0, FP32 weighted groups of `group` points added in FP64;
    group == 0, every FP32 term converted and multiply-added in FP64."""
    R, n = r.shape
    tlo, thi = p * r.min() - m, p * r.max() - m
    s2 = np.where((tlo <= 0) & (thi >= 0), 0.0, np.minimum(tlo * tlo, thi * thi))
    t = p[:, None, None] * r[None] - m[:, None, None]
    a = t * t - s2[:, None, None]
    exact = np.sum(qw[None] * np.exp2(-a), axis=(1, 2))
    e32 = f32(np.exp2(-f32(a).astype(np.float64))).astype(np.float64)
    if group == 0:
        got = np.sum(qw[None] * e32, axis=(1, 2))
    else:
        qw32 = f32(qw).astype(np.float64)
        got = np.zeros(len(p))
        for j0 in range(0, n, group):
            gs = np.zeros((len(p), R), dtype=np.float32)
            for j in range(j0, min(j0 + group, n)):
                gs = f32(qw32[None, :, j] * e32[:, :, j] + gs.astype(np.float64))
            got += np.sum(gs.astype(np.float64), axis=1)
    scale = np.exp2(-s2)
    return exact * scale, got * scale


def sight_errors(n, R, group, sights=1500, seed=20261204):
    """Relative errors per sight: stream sigma 0.1-20 kpc (the reference's
    parameter bounds), the sight's angle to the stream axis anywhere, the
    closest approach at 0.3-150 kpc (inside and beyond the region's
    0.4-106 kpc)."""
    rng = np.random.default_rng(seed + n)
    sigma = np.exp(rng.uniform(math.log(0.1), math.log(20.0), sights))
    k = np.sqrt(math.log2(math.e) / (2 * sigma ** 2))
    p = k * rng.uniform(0.05, 1.0, sights)
    m = p * np.exp(rng.uniform(math.log(0.3), math.log(150.0), sights))
    r, qw = tables(R, n)
    exact, got = column_sums(r, qw, p, m, group)
    ok = exact > 1e-300
    return got[ok] / exact[ok] - 1.0


def integral_errors(n, R, group, bundles=20, sights=300, seed=1):
    """Relative errors of integrals over bundles of sights crossing one
    stream (sigma 0.5-5 kpc, closest approach 2-60 kpc)."""
    rng = np.random.default_rng(seed + n)
    r, qw = tables(R, n)
    out = []
    for _ in range(bundles):
        sigma = math.exp(rng.uniform(math.log(0.5), math.log(5.0)))
        k = math.sqrt(math.log2(math.e) / (2 * sigma ** 2))
        rstar = math.exp(rng.uniform(math.log(2.0), math.log(60.0)))
        p = k * rng.uniform(0.3, 1.0, sights)
        m = p * rstar * rng.uniform(0.95, 1.05, sights)
        exact, got = column_sums(r, qw, p, m, group)
        out.append(np.sum(got) / np.sum(exact) - 1.0)
    return np.array(out)


@pytest.mark.parametrize("n, R, bound", [(120, 140, 1.5e-7)])
def test_group_sums_per_sight_and_bias(n, R, bound):
    """Per sight at most 1.1e-7 (against 3.8e-8 with each term in FP64, the
    weights' and products' FP32 roundings coming on top of the exponent's);
    the mean over sights within 5e-9."""
    err = sight_errors(n, R, GROUP)
    worst, mean, rms = np.max(np.abs(err)), np.mean(err), math.sqrt(np.mean(err ** 2))
    print(f"{n} points, {R} r-steps: worst {worst:.2e}, mean {mean:.2e}, rms {rms:.2e}")
    assert worst <= bound
    assert abs(mean) <= 5e-9
    assert rms <= 3e-8


@pytest.mark.parametrize("n, R", [(120, 70), (120, 140)])
def test_group_sums_in_an_integral(n, R):
    """Integrals over bundles of sights at 120 points per row: the groups
    leave at most 5e-9, of either sign (measured 3.9e-9 over 70 r-steps and
    2.0e-9 over 140; FP64 terms 2.4e-10 and 8e-11)."""
    err = integral_errors(n, R, GROUP)
    print(f"{n} points, {R} r-steps: worst {np.max(np.abs(err)):.2e}, mean {np.mean(err):.2e}")
    assert np.max(np.abs(err)) <= 5e-9


@pytest.mark.parametrize("n, R", [(8, 40), (30, 70), (64, 70)])
def test_coarse_rows_stay_on_fp64_terms(n, R):
    """Below GROUP_MIN_CONV points per row the groups would leave about
    1e-8 of an integral (measured 1.35e-8, 1.14e-8 and 9.9e-9 at 8, 30 and
    64 points: the FP32 weights' rounding, the same in every column,
    averages over few points), while FP64 terms stay near 1e-9 (1.0e-9,
    3.2e-10 and 1.9e-10)."""
    assert n < GROUP_MIN_CONV
    grouped = np.max(np.abs(integral_errors(n, R, GROUP)))
    per_term = np.max(np.abs(integral_errors(n, R, 0)))
    print(f"{n} points: grouped {grouped:.2e}, FP64 terms {per_term:.2e}")
    assert grouped > 8e-9
    assert per_term <= 1.5e-9


@pytest.mark.parametrize("n, R", [(120, 140)])
def test_whole_row_in_fp32_fails_the_bias_bound(n, R):
    """The rejected alternative, each r-step's row summed in FP32, drifts
    downwards beyond the groups' bias bound."""
    err = sight_errors(n, R, n)
    assert np.mean(err) < -5e-9
