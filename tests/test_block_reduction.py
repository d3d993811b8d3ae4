"""The resampled kernels' block reduction (block_reduce_f64 in
mw_kernels.hip), restated in numpy in the kernel's order: within each
64-lane wave an xor butterfly (lane i adds lane i ^ m for m = 32 .. 1), then
the four waves' results added in wave order.  Over positive terms its
error is at most 9 u of the exact sum (6 butterfly additions and at most 3
wave additions per term, u = 2^-53), which the kernel's comment states.
The bound is checked against exact sums (math.fsum) of positive blocks
spread over many orders of magnitude."""
import math

import numpy as np

U = 2.0 ** -53


def block_reduce(v):
    v = np.array(v, dtype=np.float64).reshape(4, 64)
    for m in (32, 16, 8, 4, 2, 1):
        idx = np.arange(64) ^ m
        v = v + v[:, idx]
    a = v[0, 0]
    for w in range(1, 4):
        a = a + v[w, 0]
    return a


def test_butterfly_lanes_agree():
    rng = np.random.default_rng(1)
    x = rng.random(256)
    v = x.reshape(4, 64).copy()
    for m in (32, 16, 8, 4, 2, 1):
        v = v + v[:, np.arange(64) ^ m]
    assert np.all(v == v[:, :1])          # every lane of a wave holds the same sum


def test_positive_block_error_within_nine_ulp():
    rng = np.random.default_rng(2)
    worst = 0.0
    for trial in range(2000):
        spread = rng.choice([0.0, 3.0, 10.0, 30.0])
        x = np.exp(rng.normal(0.0, spread, 256)) * rng.random(256)
        exact = math.fsum(x)
        err = abs(block_reduce(x) - exact) / exact
        worst = max(worst, err)
    print(f"worst relative error {worst / U:.2f} u")
    assert worst <= 9 * U


def test_sign_changing_terms_need_more():
    """With terms of both signs the error is relative to the sum of the
    magnitudes, not of the sum: the star pass's log10 P terms change sign,
    which is why it reduces in double-double."""
    x = np.array([1.0, -1.0 + 2 ** -40] * 128)
    exact = math.fsum(x)
    got = block_reduce(x)
    assert abs(got - exact) <= 9 * U * np.sum(np.abs(x))
