"""The stream-Gaussian recurrence of the resampled kernels, as an algorithm.

On a grid uniform in distance the mixed kernels take G_j = 2^-(t_j^2),
t_j = t_0 + j d, by G_{j+1} = G_j R_j, R_{j+1} = R_j C with
R_0 = 2^-(2 t_0 d + d^2), C = 2^-(2 d^2), restarting from exact values every
8 rows of 32 nodes and at any row that starts before the peak (t < 0) with
G < 2^-1000, t clamped at -40 on restart.  This restates that control flow
in numpy FP64 (same operation order) and checks it against the direct
2^-(t^2) over columns that cross the peak, start far before it (underflow
and clamping), never reach it, or pass it at the largest step the host
allows (d = 1/2): the sum over a column must match to 1e-11 relative to the
column's largest term and no inf or nan may appear.
"""
import numpy as np
import pytest

ROW = 32
RESTART_ROWS = 8


def exp2_neg(a):
    return np.exp2(-np.maximum(np.asarray(a, dtype=np.float64), -1100.0)) if np.ndim(a) else \
        float(np.exp2(-max(a, -1100.0)))


def column_by_recurrence(t_first, d, rows, w, underflow_restart=True):
    G = R = 0.0
    C = exp2_neg(2.0 * d * d)
    total = 0.0
    for k in range(rows):
        t0 = t_first + k * ROW * d
        if k % RESTART_ROWS == 0 or (underflow_restart and t0 < 0.0 and G < 2.0 ** -1000):
            tc = max(t0, -40.0)
            G = exp2_neg(tc * tc)
            R = exp2_neg(2.0 * tc * d + d * d)
        for i in range(ROW):
            total += w[k * ROW + i] * G
            G *= R
            R *= C
            assert np.isfinite(G) and np.isfinite(R)
    return total


def column_direct(t_first, d, rows, w):
    j = np.arange(rows * ROW)
    t = t_first + j * d
    return float(np.sum(w * np.exp2(-(t * t)))), float(np.max(w * np.exp2(-(t * t))))


@pytest.mark.parametrize("d", [1e-3, 0.05, 0.21, 0.5])
@pytest.mark.parametrize("t_first", [-80.0, -45.0, -38.0, -33.0, -31.0, -12.0, -1.0, 0.0, 3.0, 30.0])
def test_recurrence_matches_direct(t_first, d):
    rows = 40
    w = np.random.default_rng(7).uniform(0.5, 1.5, rows * ROW)
    got = column_by_recurrence(t_first, d, rows, w)
    want, peak = column_direct(t_first, d, rows, w)
    if peak == 0.0:
        assert got == 0.0
        return
    # relative to the column's largest term: what the kernel's 2^-e factor
    # scales to the sight's contribution
    assert abs(got - want) <= 1e-11 * max(peak, want)


def test_restart_needed_after_underflow():
    """A column that starts 35 units before the peak underflows to 0 at its
    first restart; the peak lies in its sixth row, before the periodic
    restart.  With the underflow restart the sum is exact; the same control
    flow without it loses the peak."""
    d, rows = 0.2, 40
    w = np.ones(rows * ROW)
    want, _ = column_direct(-35.0, d, rows, w)
    assert want > 1.0
    got = column_by_recurrence(-35.0, d, rows, w)
    assert abs(got - want) <= 1e-11 * want
    planted = column_by_recurrence(-35.0, d, rows, w, underflow_restart=False)
    assert abs(planted - want) > 0.5 * want
