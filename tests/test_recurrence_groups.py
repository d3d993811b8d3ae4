"""The split layout's stream recurrence, eight terms per step, pinned on the
CPU in FP64 (numpy), in the kernel's order of operations.

Along a distance grid t_j = t_0 + j d, the terms G_j = 2^-(t_j^2) obey
G_{j+1} = G_j R_j and R_{j+1} = R_j c with c = 2^-(2 d^2).  The kernel takes
eight terms per step in Horner form,
    a += G_j (w_j + R_j (w_{j+1} + R_{j+1} (... + R_{j+6} w_{j+7}))),
R_{j+k} = R_j c^k, then G_{j+8} = G_j P_j with P_j = R_j^8 c^28,
P_{j+8} = P_j c^64 and R_{j+8} = R_j c^8, restarting from the exact 2^x every
8 rows of 32 nodes.  Over that interval the sum must stay within 4e-12 of
the direct one for any start and spacing the kernel allows (d <= 1/2,
t_0 >= -40); a start P_0 of R^8 c^24 (one factor c^4 short) must fail.
"""
import numpy as np

ROWS, N = 8, 32


def grouped(w, t0, d, short_start=False):
    c = 2.0 ** (-2 * d * d)
    c4 = (c * c) * (c * c)
    c8 = c4 * c4
    c16 = c8 * c8
    c24 = c16 * c8
    c64 = (c16 * c16) * (c16 * c16)
    tc = max(t0, -40.0)
    G = 2.0 ** (-tc * tc)
    R = 2.0 ** (-(2 * tc * d + d * d))
    q = R * R * c
    q2 = q * q
    P = (q2 * q2) * (c16 * c4 if short_start else c24)
    a = 0.0
    for i in range(0, w.size, 8):
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
    return a


def worst_error(**kw):
    rng = np.random.default_rng(20261204)
    worst = 0.0
    for _ in range(1500):
        d = rng.uniform(0.01, 0.5)
        t0 = rng.uniform(-30.0, 5.0)
        w = rng.uniform(0.1, 10.0, ROWS * N)
        t = t0 + d * np.arange(ROWS * N)
        exact = np.sum(w * np.exp2(-t * t))
        if exact > 1e-280:
            worst = max(worst, abs(grouped(w, t0, d, **kw) / exact - 1.0))
    return worst


def test_grouped_recurrence_bound():
    worst = worst_error()
    print(f"grouped recurrence over {ROWS} rows: worst {worst:.2e}")
    assert worst <= 4e-12


def test_grouped_recurrence_catches_a_short_start():
    assert worst_error(short_start=True) > 1e-3
