"""The resampled quadrature's weight folding (mw_resample_weights), on the CPU.

Pins: polynomial exactness up to the stencil degree (what Lagrange folding
guarantees by construction), conservation of the total weight, and the
accuracy on the integrand's narrowest profiles over a full-size point set
(700 r-steps x 120 convolution points): a stream Gaussian of 1 kpc crossing
at up to 100 kpc and a background peak 0.14 kpc from the Galactic centre, at
the grid spacings the level rule picks for them.
"""
import numpy as np
import pytest

from milkyway_client_amd import separation as sep
from oracle import oracle as O


@pytest.fixture(scope="module")
def points():
    R, n = 700, 120
    t = O.r_constants(n, R, 16.0, 6.5 / R, 1.0, 1, 0.0, 1.0)
    rp = np.asarray(t["r_point"]).reshape(R, n)
    w = (np.asarray(t["irv"]) * np.asarray(t["reff_xr_rp3"]))[:, None] * np.asarray(t["qw_r3_N"]).reshape(R, n)
    g = 5 * (np.log10(rp * 1000) - 1) + 4.2
    return rp.ravel(), g.ravel(), w.ravel()


def r_of(g):
    return 10 ** ((g - 4.2) / 5 + 1) / 1000


def test_polynomials_are_exact(points):
    _, g, w = points
    for delta in (0.02, 0.005):
        lo, K = sep.resample_weights(g, w, delta)
        gm = lo + delta * np.arange(K.size)
        c = 19.0   # moments about the centre of the range keep the sums well conditioned
        for j in range(16):
            want = np.sum(w * (g - c) ** j)
            got = np.sum(K * (gm - c) ** j)
            assert abs(got - want) <= 1e-11 * np.sum(np.abs(w * (g - c) ** j)), (delta, j)


def test_distance_grid_polynomials_are_exact(points):
    rp, _, w = points
    lo, K = sep.resample_weights(rp, w, 0.25)
    rm = lo + 0.25 * np.arange(K.size)
    c = 50.0
    for j in range(16):
        want = np.sum(w * (rp - c) ** j)
        got = np.sum(K * (rm - c) ** j)
        assert abs(got - want) <= 1e-11 * np.sum(np.abs(w * (rp - c) ** j)), j


def test_distance_grid_narrow_stream(points):
    """A grid uniform in distance at 0.2 sigma: the stream Gaussian's width in
    distance is at least sigma at any distance."""
    rp, _, w = points
    lo, K = sep.resample_weights(rp, w, 0.2)
    rm = lo + 0.2 * np.arange(K.size)
    rng = np.random.default_rng(4)
    k2 = np.log2(np.e) / 2.0           # sigma = 1 kpc
    for _ in range(300):
        rc, pu, e = rng.uniform(2, 100), rng.uniform(0.3, 1.0), rng.uniform(0, 3)
        f = lambda r: np.exp2(-k2 * ((pu * (r - rc)) ** 2 + e))
        exact = np.sum(w * f(rp))
        assert abs(np.sum(K * f(rm)) - exact) <= 1e-11 * exact


def test_grid_covers_points_with_margin(points):
    _, g, w = points
    lo, K = sep.resample_weights(g, w, 0.01)
    assert lo < g.min() - 7 * 0.01
    assert lo + 0.01 * (K.size - 1) > g.max() + 7 * 0.01
    assert np.count_nonzero(K) > 0.9 * K.size


@pytest.mark.parametrize("delta,tol", [(0.005, 1e-11), (0.0025, 1e-12)])
def test_narrow_stream_accuracy(points, delta, tol):
    rp, g, w = points
    lo, K = sep.resample_weights(g, w, delta)
    rm = r_of(lo + delta * np.arange(K.size))
    rng = np.random.default_rng(1)
    k2 = np.log2(np.e) / 2.0           # sigma = 1 kpc
    for _ in range(300):
        rc, pu, e = rng.uniform(2, 100), rng.uniform(0.3, 1.0), rng.uniform(0, 3)
        f = lambda r: np.exp2(-k2 * ((pu * (r - rc)) ** 2 + e))
        exact = np.sum(w * f(rp))
        assert abs(np.sum(K * f(rm)) - exact) <= tol * exact


def test_background_near_galactic_centre(points):
    rp, g, w = points
    lo, K = sep.resample_weights(g, w, 0.005)
    rm = r_of(lo + 0.005 * np.arange(K.size))
    rng = np.random.default_rng(2)
    for _ in range(100):
        u = np.array([1.0, 0, 0]) + rng.normal(size=3) * 0.1
        u /= np.linalg.norm(u)
        q, r0 = rng.uniform(0.3, 1.0), rng.uniform(4, 8)

        def f(r):
            rg = np.sqrt((r * u[0] - 8.5) ** 2 + (r * u[1]) ** 2 + (r * u[2] / q) ** 2)
            return 1 / (rg * (rg + r0) ** 3)

        exact = np.sum(w * f(rp))
        assert abs(np.sum(K * f(rm)) - exact) <= 1e-10 * exact


def test_bad_input():
    with pytest.raises(ValueError):
        sep.resample_weights([1.0, 2.0], [1.0, 1.0], 0.0)
