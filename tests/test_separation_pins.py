"""Pins of the oracle's stream separation (separation.c:37-254), on the CPU.

Each check compares the oracle with something it does not share code with:
the independent numpy model (tests/pin_math.py: rotation-matrix geometry,
Rodrigues' rotation, the posterior written from the mixture), glibc's own
srand48/drand48 through ctypes, and inverse-CDF sampling with
numpy.searchsorted.  Comparisons with the numpy model use at most 10
convolution points: the reference's Gauss-Legendre weights carry a ~1e-9
relative error from 30 points up, which the model (numpy's leggauss) does
not share.
"""
import ctypes as C

import numpy as np
import pytest

import pin_math as PM
from helpers import fixture
from milkyway_client_amd import synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder
from oracle import oracle as O

WEDGES = [9, 10, 11, 12, 20, 21, 46, 47, 79, 82, 86]


@pytest.mark.parametrize("sgr", [0, 1])
def test_stripe_normal_is_the_stripe_pole(sgr):
    """SDSS: stripe_normal is the nu = 90 pole of the stripe's great circle.
    Sagittarius: sgr_stripe_normal is the unit normal (lambda, 0) x (lambda, 90)."""
    for w in WEDGES:
        n = O.stripe_normal(w, sgr)
        ref = PM.stripe_normal(w, sgr)
        assert abs(np.linalg.norm(n) - 1) < 1e-14
        assert np.max(np.abs(n - ref)) < 1e-12, w


@pytest.mark.parametrize("sgr", [0, 1])
def test_stripe_normal_is_perpendicular_to_the_stripe(sgr):
    """Every point of the stripe (nu = 0) lies in the plane the normal
    defines, through the oracle's own survey chain (a different path)."""
    for w in WEDGES:
        n = O.stripe_normal(w, sgr)
        for mu in np.linspace(0, 350, 15):
            l, b = O.gc_to_lb(w, sgr, mu, 0.0)
            assert abs(n @ PM.unit_los(l, b)) < 1e-10, (w, mu)


def test_get_transform_is_the_rotation_to_z():
    rng = np.random.default_rng(5)
    z = np.array([0.0, 0.0, 1.0])
    for _ in range(50):
        f = rng.normal(size=3)
        f /= np.linalg.norm(f)
        M = O.get_transform(f, z)
        assert np.max(np.abs(M @ M.T - np.eye(3))) < 1e-14
        assert abs(np.linalg.det(M) - 1) < 1e-14
        assert np.max(np.abs(M @ f - z)) < 1e-14
        assert np.max(np.abs(M - PM.rotation_to_z(f))) < 1e-13


def test_transform_point_closed_form():
    """xyz_mag + the xsun shift + the rotation collapse to rot (g u) with
    g = 5 log10(100 r) - 6.8 along the star's own direction."""
    rng = np.random.default_rng(6)
    M = O.get_transform(O.stripe_normal(86), [0.0, 0.0, 1.0])
    stars = np.stack([rng.uniform(0, 360, 200), rng.uniform(-80, 80, 200), rng.uniform(0.5, 80, 200)], 1)
    ref = PM.separation_positions(stars, M)
    for s, want in zip(stars, ref):
        got = O.transform_point(O.lbr2xyz(*s), M)
        assert np.max(np.abs(got - want)) < 1e-11


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_prob_ok_is_inverse_cdf_sampling(n):
    rng = np.random.default_rng(10 + n)
    for _ in range(2000):
        p = rng.dirichlet(np.ones(n + 1))[:n]
        r = rng.uniform()
        cum = np.cumsum(p)
        want = 0 if r > cum[-1] else int(np.searchsorted(cum, r, side="left")) + 1
        assert O.prob_ok(p, r) == want
    assert O.prob_ok(np.ones(n) * 0.1, 0.95) == 0


def test_prob_ok_edges():
    assert O.prob_ok([0.3, 0.2], 0.3) == -1        # r == p[0]: unset in the reference
    assert O.prob_ok([0.3, 0.2, 0.1, 0.1], 0.3) == 1
    assert O.prob_ok([], 0.5) == -2 and O.prob_ok([0.1] * 5, 0.5) == -2


def glibc_drand48(seed: int, count: int) -> np.ndarray:
    libc = C.CDLL("libc.so.6")
    libc.srand48.argtypes = [C.c_long]
    libc.drand48.restype = C.c_double
    libc.srand48(seed)
    return np.array([libc.drand48() for _ in range(count)])


def run(p, stars, seed, I=None):
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    if I is None:
        ref = O.evaluate_full(ph, sh)
        I = (ref.background_integral, ref.stream_integrals)
    return O.separation(ph, sh, I[0], I[1], seed), I


@pytest.mark.parametrize("fid,n_stars", [("86", 2000), ("11", 600)])
def test_separation_against_numpy_model_and_glibc(fid, n_stars):
    p, stars = fixture(fid, n_stars=n_stars, convolve=10)
    seed = 20240917
    sep, (I_bg, I_st) = run(p, stars, seed)
    S = len(p.streams)
    want = PM.membership(p, stars, I_bg, I_st)
    assert np.max(np.abs(sep.probability - want)) < 1e-12
    # the reference's running sum and counts
    assert np.allclose(sep.expected, sep.probability.sum(axis=0), rtol=1e-12, atol=0)
    assert list(sep.separated) == [int(np.sum(sep.assignment == s + 1)) for s in range(S)]
    # the i-th star draws glibc's i-th drand48() after srand48(seed)
    r = glibc_drand48(seed, len(stars))
    cum = np.cumsum(sep.probability, axis=1)
    want_a = np.where(r > cum[:, -1], 0, np.argmax(cum >= r[:, None], axis=1) + 1)
    assert np.array_equal(sep.assignment, want_a)
    M = PM.rotation_to_z(PM.stripe_normal(p.wedge, p.sgr_coordinates))
    assert np.max(np.abs(sep.xyz - PM.separation_positions(stars, M))) < 1e-11


def test_background_weight_does_not_enter_separation():
    """separation.c:127-140 weights the background with exp(0) = 1 whatever
    background_weight holds (the likelihood uses exp(background_weight))."""
    p, stars = fixture("79", n_stars=300, convolve=10)
    p.background_weight = 0.7
    sep, (I_bg, I_st) = run(p, stars, 1)
    assert np.max(np.abs(sep.probability - PM.membership(p, stars, I_bg, I_st))) < 1e-12


def test_separation_sgr_wedge_multi_stream():
    p, stars = synth.workload(7, synth.WedgeShape(16, 24, 12, 10, 400), 4)
    p.sgr_coordinates = 1
    sep, (I_bg, I_st) = run(p, stars, 77)
    assert np.max(np.abs(sep.probability - PM.membership(p, stars, I_bg, I_st))) < 1e-12
    M = PM.rotation_to_z(PM.stripe_normal(p.wedge, 1))
    assert np.max(np.abs(sep.xyz - PM.separation_positions(stars, M))) < 1e-11


def test_assignment_frequencies_follow_probabilities():
    """Law of large numbers: counts per stream track the summed probabilities."""
    p, stars = fixture("86", n_stars=2000, convolve=20)
    sep, _ = run(p, stars, 3)
    for s in range(len(p.streams)):
        e = sep.expected[s]
        sd = np.sqrt(np.sum(sep.probability[:, s] * (1 - sep.probability[:, s])))
        assert abs(sep.separated[s] - e) < 5 * sd + 1
