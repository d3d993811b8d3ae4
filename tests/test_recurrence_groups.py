"""The split layout's stream recurrence, eight terms per step, pinned on the
CPU in FP64 (numpy), in the kernel's order of operations.

Along a distance grid t_j = t_0 + j d, the terms are carried relative to the
sight's largest on-grid term, G_j = 2^-(t_j^2 - s^2) with s^2 the smallest
t^2 over the grid's nodes, and obey G_{j+1} = G_j R_j and R_{j+1} = R_j c
with c = 2^-(2 d^2).  The kernel takes eight terms per step in Horner form,
    a += G_j (w_j + R_j (w_{j+1} + R_{j+1} (... + R_{j+6} w_{j+7}))),
R_{j+k} = R_j c^k, then G_{j+8} = G_j P_j with P_j = R_j^8 c^28,
P_{j+8} = P_j c^64 and R_{j+8} = R_j c^8, restarting from the exact 2^x every
8 rows of 32 nodes and at a row that starts before the peak with G < 2^-1000
(t clamped at -sqrt(s^2 + 1600) on restart), G and the sum carried scaled
by 2^B (B = ceil(t^2 - s^2 - 1000) at a restart before the peak, short of
the clamp, more than 2^-1000 below the largest term, else 0), and running no terms for a sight with s^2 >= 1100.  Over that interval the shifted sum must stay
within 4e-12 of the direct one for any start and spacing the kernel allows
(d <= 1/2, t_0 from -60 to 5: the Gaussian's peak ahead of the interval, on
it, or behind it; starts in subnormal range and beyond the clamp).  A start P_0 of R^8 c^24 (one factor c^4 short) must
fail, and so must the same steps without the shift (terms below 2^-1022
start subnormal or at 0), without the scale, or without the restart on
underflow.
"""
import numpy as np

ROWS, N = 8, 32
GONE_T2 = 1100.0


def min_t2(t_first, t_last):
    if t_first <= 0.0 <= t_last:
        return 0.0
    t = min(abs(t_first), abs(t_last))
    return t * t


def grouped(w, t0, d, s2, short_start=False, underflow_restart=True, scale=True):
    """The shifted sum over the interval's nodes, s2 the shift."""
    if s2 >= GONE_T2:
        return 0.0
    c = 2.0 ** (-2 * d * d)
    c4 = (c * c) * (c * c)
    c8 = c4 * c4
    c16 = c8 * c8
    c24 = c16 * c8
    c64 = (c16 * c16) * (c16 * c16)
    a = 0.0
    G = R = P = 0.0
    B = 0
    for i in range(0, w.size, 8):
        t_row = t0 + i * d
        if i == 0 or (underflow_restart and i % N == 0 and t_row < 0.0 and G < 2.0 ** -1000):
            # restart: at the interval's first row, and at a row that starts
            # before the peak with G < 2^-1000
            lim = np.sqrt(s2 + 1600.0)
            tc = max(t_row, -lim)
            R = 2.0 ** (-(2 * tc * d + d * d))
            D = tc * tc - s2
            deep = scale and -lim < t_row < 0.0 and D > 1000.0
            Bn = int(np.ceil(D - 1000.0)) if deep else 0
            a = float(np.ldexp(a, Bn - B))
            B = Bn
            G = 2.0 ** max(-(D - Bn), -1100.0)   # a clamped start: 2^-1600 held at 2^-1100 = 0
            q = R * R * c
            q2 = q * q
            P = (q2 * q2) * (c16 * c4 if short_start else c24)
        r = [R]
        for _ in range(6):
            r.append(r[-1] * c)
        x = r[6] * w[i + 7] + w[i + 6]
        for k in range(5, -1, -1):
            x = r[k] * x + w[i + k]
        a += G * x
        G *= P
        P *= c64
        R *= c8
    return float(np.ldexp(a, -B))


def worst_error(shift=True, **kw):
    rng = np.random.default_rng(20261204)
    worst = 0.0
    gone = 0
    for _ in range(1500):
        d = rng.uniform(0.01, 0.5)
        t0 = rng.uniform(-60.0, 5.0)
        w = rng.uniform(0.1, 10.0, ROWS * N)
        t = t0 + d * np.arange(ROWS * N)
        s2 = min_t2(t[0], t[-1])
        exact = np.sum(w * np.exp2(-(t * t - s2)))
        if s2 >= GONE_T2:
            # every term below 2^-1100 of its weight: not run
            assert grouped(w, t0, d, s2, **kw) == 0.0
            gone += 1
            continue
        if shift:
            got = grouped(w, t0, d, s2, **kw)
        else:
            # the unshifted sum, rescaled by 2^s2 for comparison (in two
            # factors: 2^1100 overflows)
            half = 2.0 ** (0.5 * s2)
            got = grouped(w, t0, d, 0.0, **kw) * half * half
        worst = max(worst, abs(got / exact - 1.0))
    return worst, gone


def test_grouped_recurrence_bound():
    worst, gone = worst_error()
    print(f"grouped recurrence over {ROWS} rows: worst {worst:.2e} ({gone} intervals not run)")
    assert worst <= 4e-12
    assert gone > 0


def test_grouped_recurrence_catches_a_short_start():
    assert worst_error(short_start=True)[0] > 1e-3


def test_grouped_recurrence_needs_the_shift():
    assert worst_error(shift=False)[0] > 1e-3


def test_grouped_recurrence_needs_the_underflow_restart():
    assert worst_error(underflow_restart=False)[0] > 1e-3


def test_grouped_recurrence_needs_the_scale():
    assert worst_error(scale=False)[0] > 1e-3
