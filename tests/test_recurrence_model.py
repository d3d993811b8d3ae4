"""The stream-Gaussian recurrence of the resampled kernels, as an algorithm.

On a grid uniform in distance the mixed kernels take G_j = 2^-(t_j^2 - s^2),
t_j = t_0 + j d, by G_{j+1} = G_j R_j, R_{j+1} = R_j C with
R_0 = 2^-(2 t_0 d + d^2), C = 2^-(2 d^2).  s^2 is the smallest t^2 over the
column's nodes (0 when the Gaussian's peak lies on the grid): the terms are
carried relative to the column's largest, and the column's sum is scaled by
2^-s^2 at the end.  The recurrence restarts from exact values every 8 rows of
32 nodes and at any row that starts before the peak (t < 0) with
G < 2^-1000, t clamped at -sqrt(s^2 + 1600) on restart (G = 2^-1600 = 0 when
clamped);
G and the column's running sum are carried scaled by 2^B, with
B = ceil(t_0^2 - s^2 - 1000) set at a restart whose start lies before the
peak, short of the clamp, and more than 2^-1000 below the column's largest
term (else 0), so that G never starts
subnormal; a column with s^2 >= 1100 is not run (its sum is below 2^-1100 of
its weights); a row whose terms all lie below 2^-60 of the column's largest
is skipped, with G set to 0, when that holds on every lane of the wave.  This restates that control flow
in numpy FP64 (same operation order) and checks it against the direct
2^-(t^2 - s^2) over columns that cross the peak, start far before it
(underflow and clamping), never reach it, stop short of it by up to 33 units
(s^2 near 1100, where unshifted terms are subnormal), or pass it at the
largest step the host allows (d = 1/2): the shifted sum over a column must
match to 1e-11 relative to the column's largest term and no inf or nan may
appear.  The same control flow without the shift, without the scaling,
without the underflow restart, or without the reset on a skipped row, must
fail.
"""
import numpy as np
import pytest

ROW = 32
RESTART_ROWS = 8
NEGLIGIBLE_LOG2 = 60.0
RESTART_CLAMP_T2 = 1600.0
GONE_T2 = 1100.0
SCALE_T2 = 1000.0


def exp2_neg(a):
    """2^-a with the exponent held at -1100 and above, as the kernel's does
    (so a large a gives 0)."""
    return np.exp2(np.maximum(-np.asarray(a, dtype=np.float64), -1100.0)) if np.ndim(a) else \
        float(np.exp2(max(-a, -1100.0)))


def min_t2(t_first, t_last):
    if t_first <= 0.0 <= t_last:
        return 0.0
    t = min(abs(t_first), abs(t_last))
    return t * t


def column_s2(t_first, d, rows):
    return min_t2(t_first, t_first + (rows * ROW - 1) * d)


def columns_by_recurrence(t_first, d, rows, w, underflow_restart=True, skip=True, zero_on_skip=True,
                          shift=True, scale=True):
    """The kernel's control flow over a wave of lanes (one column each,
    t_first per lane): the skip is decided for the whole wave, restarts per
    lane.  Returns each lane's sum relative to 2^-s^2 (its shifted sum) and
    s^2; shift=False carries the unshifted terms (s^2 taken as 0 inside the
    recurrence) and rescales the result by 2^s^2 for comparison."""
    t_first = np.atleast_1d(np.asarray(t_first, dtype=np.float64))
    lanes = t_first.size
    G = np.zeros(lanes)
    R = np.zeros(lanes)
    C = exp2_neg(2.0 * d * d)
    s2 = np.array([column_s2(t, d, rows) for t in t_first])
    s2_run = s2 if shift else np.zeros(lanes)
    total = np.zeros(lanes)
    B = np.zeros(lanes, dtype=np.int64)
    for k in range(rows):
        t0 = t_first + k * ROW * d
        t1 = t0 + (ROW - 1) * d
        if skip and all(min_t2(a, b) > f + NEGLIGIBLE_LOG2 for a, b, f in zip(t0, t1, s2_run)):
            if zero_on_skip:
                G[:] = 0.0
            continue
        for j in range(lanes):
            if k % RESTART_ROWS == 0 or (underflow_restart and t0[j] < 0.0 and G[j] < 2.0 ** -1000):
                if s2_run[j] >= GONE_T2:
                    G[j] = R[j] = 0.0
                    continue
                lim = np.sqrt(s2_run[j] + RESTART_CLAMP_T2)
                tc = max(t0[j], -lim)
                R[j] = exp2_neg(2.0 * tc * d + d * d)
                D = tc * tc - s2_run[j]
                # scaled only before the peak and short of the clamp; a
                # clamped start keeps B = 0 and its G = 2^-1600 is 0
                deep = scale and -lim < t0[j] < 0.0 and D > SCALE_T2
                Bn = int(np.ceil(D - SCALE_T2)) if deep else 0
                total[j] = np.ldexp(total[j], Bn - B[j])
                B[j] = Bn
                G[j] = exp2_neg(D - Bn)
        for i in range(ROW):
            total += w[k * ROW + i] * G
            G *= R
            R *= C
            assert np.all(np.isfinite(G)) and np.all(np.isfinite(R)) and np.all(np.isfinite(total))
    total = np.ldexp(total, -B)
    if not shift:
        half = np.exp2(0.5 * s2)   # 2^s2 in two factors: 2^1100 would overflow
        total = total * half * half
    return total, s2


def column_by_recurrence(t_first, d, rows, w, **kw):
    got, s2 = columns_by_recurrence([t_first], d, rows, w, **kw)
    return float(got[0]), float(s2[0])


def column_direct(t_first, d, rows, w):
    """The column's shifted sum and largest shifted term, directly."""
    j = np.arange(rows * ROW)
    t = t_first + j * d
    s2 = column_s2(t_first, d, rows)
    terms = w * np.exp2(-(t * t - s2))
    return float(np.sum(terms)), float(np.max(terms))


def check_column(t_first, d, rows, w):
    got, s2 = column_by_recurrence(t_first, d, rows, w)
    want, peak = column_direct(t_first, d, rows, w)
    if s2 >= GONE_T2:
        # not run: every term is below 2^-1100 of its weight
        assert got == 0.0
        assert np.log2(want) - s2 < -1090.0
        return
    # relative to the column's largest term: what the kernel's 2^-(e + s^2)
    # factor scales to the sight's contribution
    assert abs(got - want) <= 1e-11 * max(peak, want)


@pytest.mark.parametrize("d", [1e-3, 0.05, 0.21, 0.5])
@pytest.mark.parametrize("t_first", [-80.0, -45.0, -38.0, -36.0, -33.0, -31.0, -12.0, -1.0, 0.0, 3.0,
                                     30.0, 33.0, 34.0])
def test_recurrence_matches_direct(t_first, d):
    rows = 40
    w = np.random.default_rng(7).uniform(0.5, 1.5, rows * ROW)
    check_column(t_first, d, rows, w)


@pytest.mark.parametrize("t_last", [-1.0, -20.0, -31.6, -32.5, -33.1])
def test_peak_beyond_the_grid(t_last):
    """Columns that stop short of the peak, their largest term 2^-(t_last^2)
    down to 2^-1096 (subnormal, or 0, without the shift).  The shifted sum
    stays within 1e-11 of its largest term; the same control flow without
    the shift fails once the largest term is subnormal."""
    d, rows = 0.01, 12
    t_first = t_last - (rows * ROW - 1) * d
    w = np.random.default_rng(3).uniform(0.5, 1.5, rows * ROW)
    check_column(t_first, d, rows, w)
    want, _ = column_direct(t_first, d, rows, w)
    planted, _ = column_by_recurrence(t_first, d, rows, w, shift=False)
    if t_last * t_last > 1030.0:
        assert abs(planted - want) > 1e-6 * want


def test_restart_needed_after_underflow():
    """A column that starts 35 units before the peak underflows to 0 at its
    first restart; the peak lies in its sixth row, before the periodic
    restart.  With the underflow restart the sum is exact; the same control
    flow without it loses the peak."""
    d, rows = 0.2, 40
    w = np.ones(rows * ROW)
    want, _ = column_direct(-35.0, d, rows, w)
    assert want > 1.0
    got, _ = column_by_recurrence(-35.0, d, rows, w)
    assert abs(got - want) <= 1e-11 * want
    planted, _ = column_by_recurrence(-35.0, d, rows, w, underflow_restart=False)
    assert abs(planted - want) > 0.5 * want


@pytest.mark.parametrize("t_first", [-60.0, -20.0, -9.0, 0.5, 12.0])
def test_row_skipping_drops_only_negligible_terms(t_first):
    """With skipping, the sum stays within 1e-11 of the direct one for a
    column that spans the peak and its far tails."""
    d, rows = 0.25, 40
    w = np.random.default_rng(11).uniform(0.5, 1.5, rows * ROW)
    check_column(t_first, d, rows, w)


def test_rows_are_skipped():
    """A column from -60 through the peak to +260 holds rows whose every term
    is below 2^-60 of the peak; the skip takes them (the sum agrees with the
    run without skipping to 1e-12: the skipped terms are below 1e-16 of it,
    and a restart after a skip rounds differently from the carried state)."""
    d, rows = 0.25, 40
    w = np.random.default_rng(11).uniform(0.5, 1.5, rows * ROW)
    t0 = -60.0 + ROW * d * np.arange(rows)
    skipped = sum(min_t2(a, a + (ROW - 1) * d) > NEGLIGIBLE_LOG2 for a in t0)
    assert skipped >= rows // 2
    with_skip, _ = column_by_recurrence(-60.0, d, rows, w)
    without, _ = column_by_recurrence(-60.0, d, rows, w, skip=False)
    assert abs(with_skip - without) <= 1e-12 * without


def test_skipped_row_must_reset_the_state():
    """Two lanes of one wave: lane A peaks in row 0, lane B in row 3, and row
    2 is negligible for both, so the wave skips it.  B's state then stands
    one row behind; setting G to 0 makes B restart at row 3.  The same
    control flow without the reset continues B from the wrong positions."""
    d, rows = 0.25, 12
    w = np.random.default_rng(5).uniform(0.5, 1.5, rows * ROW)
    lanes = [-4.0, -36.0]
    want = np.array([column_direct(t, d, rows, w)[0] for t in lanes])
    got, s2 = columns_by_recurrence(lanes, d, rows, w)
    assert np.all(s2 == 0.0)
    assert np.all(np.abs(got - want) <= 1e-11 * want)
    planted, _ = columns_by_recurrence(lanes, d, rows, w, zero_on_skip=False)
    assert abs(planted[1] - want[1]) > 1e-3 * want[1]


def test_deep_start_needs_the_scale():
    """Two lanes of one wave: lane A peaks in row 0, so the wave runs from
    row 0; lane B's row 0 starts 2^-1073.6 below its peak (t = -32.77,
    d = 0.342), a subnormal G with about one significant bit, and grows to
    2^-488 within the row, so no underflow restart follows before its peak
    in row 3.  With the scale B starts G at 2^-1000 with full precision and
    both sums match; without it, B's sum is off by more than 1 %."""
    d, rows = 0.342, 12
    w = np.random.default_rng(9).uniform(0.5, 1.5, rows * ROW)
    lanes = [-4.0, -32.77]
    want = np.array([column_direct(t, d, rows, w)[0] for t in lanes])
    got, s2 = columns_by_recurrence(lanes, d, rows, w)
    assert np.all(s2 == 0.0)
    assert np.all(np.abs(got - want) <= 1e-11 * want)
    planted, _ = columns_by_recurrence(lanes, d, rows, w, scale=False)
    assert abs(planted[1] - want[1]) > 1e-2 * want[1]


@pytest.mark.parametrize("t_b", [-32.4, -32.6, -32.77, -36.0, -38.5, -45.0, -52.0])
@pytest.mark.parametrize("d", [0.12, 0.342, 0.5])
def test_lane_behind_the_wave(t_b, d):
    """A lane B that starts well before its peak in a wave that runs every
    row from row 0 (lane A peaks there): B's sum stays within 1e-11 of the
    direct one for starts across the subnormal range and the clamp."""
    rows = 16
    w = np.random.default_rng(13).uniform(0.5, 1.5, rows * ROW)
    lanes = [-2.0, t_b]
    want = np.array([column_direct(t, d, rows, w) for t in lanes])
    got, s2 = columns_by_recurrence(lanes, d, rows, w)
    assert np.all(np.abs(got - want[:, 0]) <= 1e-11 * np.maximum(want[:, 0], want[:, 1]))
