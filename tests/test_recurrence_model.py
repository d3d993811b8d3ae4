"""The stream-Gaussian recurrence of the resampled kernels, as an algorithm.

On a grid uniform in distance the mixed kernels take G_j = 2^-(t_j^2),
t_j = t_0 + j d, by G_{j+1} = G_j R_j, R_{j+1} = R_j C with
R_0 = 2^-(2 t_0 d + d^2), C = 2^-(2 d^2), restarting from exact values every
8 rows of 32 nodes and at any row that starts before the peak (t < 0) with
G < 2^-1000, t clamped at -40 on restart, and skipping (with G set to 0) a
row whose terms all lie below 2^-70 of the column's largest.  This restates that control flow
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


NEGLIGIBLE_LOG2 = 70.0


def min_t2(t_first, t_last):
    if t_first <= 0.0 <= t_last:
        return 0.0
    t = min(abs(t_first), abs(t_last))
    return t * t


def columns_by_recurrence(t_first, d, rows, w, underflow_restart=True, skip=True, zero_on_skip=True):
    """The kernel's control flow over a wave of lanes (one column each,
    t_first per lane): the skip is decided for the whole wave, restarts per
    lane.  Returns each lane's sum."""
    t_first = np.atleast_1d(np.asarray(t_first, dtype=np.float64))
    lanes = t_first.size
    G = np.zeros(lanes)
    R = np.zeros(lanes)
    C = exp2_neg(2.0 * d * d)
    floor_t2 = np.array([min_t2(t, t + (rows * ROW - 1) * d) for t in t_first]) + NEGLIGIBLE_LOG2
    total = np.zeros(lanes)
    for k in range(rows):
        t0 = t_first + k * ROW * d
        t1 = t0 + (ROW - 1) * d
        if skip and all(min_t2(a, b) > f for a, b, f in zip(t0, t1, floor_t2)):
            if zero_on_skip:
                G[:] = 0.0
            continue
        for j in range(lanes):
            if k % RESTART_ROWS == 0 or (underflow_restart and t0[j] < 0.0 and G[j] < 2.0 ** -1000):
                tc = max(t0[j], -40.0)
                G[j] = exp2_neg(tc * tc)
                R[j] = exp2_neg(2.0 * tc * d + d * d)
        for i in range(ROW):
            total += w[k * ROW + i] * G
            G *= R
            R *= C
            assert np.all(np.isfinite(G)) and np.all(np.isfinite(R))
    return total


def column_by_recurrence(t_first, d, rows, w, **kw):
    return float(columns_by_recurrence([t_first], d, rows, w, **kw)[0])


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


@pytest.mark.parametrize("t_first", [-60.0, -20.0, -9.0, 0.5, 12.0])
def test_row_skipping_drops_only_negligible_terms(t_first):
    """With skipping, the sum stays within 1e-11 of the direct one, and rows
    really are skipped for a column that spans the peak and its far tails."""
    d, rows = 0.25, 40
    w = np.random.default_rng(11).uniform(0.5, 1.5, rows * ROW)
    want, peak = column_direct(t_first, d, rows, w)
    got = column_by_recurrence(t_first, d, rows, w)
    if peak == 0.0:
        assert got == 0.0
        return
    assert abs(got - want) <= 1e-11 * max(peak, want)


def test_skipped_row_must_reset_the_state():
    """Two lanes of one wave: lane A peaks in row 0, lane B in row 3, and row
    2 is negligible for both, so the wave skips it.  B's state then stands
    one row behind; setting G to 0 makes B restart at row 3.  The same
    control flow without the reset continues B from the wrong positions."""
    d, rows = 0.25, 12
    w = np.random.default_rng(5).uniform(0.5, 1.5, rows * ROW)
    lanes = [-4.0, -36.0]
    want = np.array([column_direct(t, d, rows, w)[0] for t in lanes])
    got = columns_by_recurrence(lanes, d, rows, w)
    assert np.all(np.abs(got - want) <= 1e-11 * want)
    planted = columns_by_recurrence(lanes, d, rows, w, zero_on_skip=False)
    assert abs(planted[1] - want[1]) > 1e-3 * want[1]
