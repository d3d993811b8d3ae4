"""Pins the CPU oracle (the checker of the HIP library) to what the
mathematics fixes, since the reference ships no golden outputs and its app
cannot be built here.

Each check is chosen so that a plausible transcription mistake in the oracle
(a dropped term, a wrong sign or index, a swapped operand) fails it:
closed forms, invariants (telescoping sums, the Gaussian convolution
integral, great-circle planes), special cases (stream axis through the line
of sight, disabled streams, q == 0, zero-probability stars), and an
independent numpy model of the whole evaluation written from the
mathematics (tests/pin_math.py).
"""
import math

import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

import pin_math as PM
from helpers import ctest_vectors, fixture, holders, rel
from milkyway_client_amd import params_io, synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder
from oracle import oracle as O

ERF3 = math.erf(3 / math.sqrt(2))


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


# --------------------------------------------------------------------------
# Gauss-Legendre quadrature (numericalIntegration.c:67-104)
# --------------------------------------------------------------------------

def test_gauss_legendre_two_point_closed_form():
    x, w = O.gauss_legendre(2)
    assert np.allclose(x, [-1 / math.sqrt(3), 1 / math.sqrt(3)], rtol=0, atol=1e-15)
    assert np.allclose(w, [1.0, 1.0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [3, 5, 10, 30, 60, 120])
def test_gauss_legendre_nodes_and_weights(n):
    x, w = O.gauss_legendre(n)
    X, W = leggauss(n)
    assert np.all(np.diff(x) > 0), "nodes ascend from -1 to 1"
    assert np.max(np.abs(x - X)) <= 5e-15
    # The reference evaluates P'_n at the last-but-one Newton iterate, so its
    # weights carry up to ~1e-9 relative error (tolerance EPS = 3e-11).
    assert np.max(np.abs(w - W) / W) <= 2e-9
    assert abs(w.sum() - 2.0) <= 1e-10
    assert np.allclose(w, w[::-1], rtol=0, atol=0), "weights are symmetric"


def test_gauss_legendre_polynomial_exactness():
    n = 10
    x, w = O.gauss_legendre(n)
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) <= 1e-11


@pytest.mark.parametrize("n", [30, 60, 120])
def test_convolution_weights_integrate_the_gaussian(n):
    """sum_i qw_i / r_i^3 * 3 sigma = sum_i W_i N(dx_i) * 3 sigma = erf(3 / sqrt 2):
    the magnitude-error Gaussian integrated over +-3 sigma."""
    t = O.r_constants(n, 4, 16.0, 1.6, 0.6, 2, -1.25, 1.25)
    s = np.sum(t["qw_r3_N"] / t["r_point"] ** 3, axis=1) * 3 * 0.6
    assert np.max(np.abs(s - ERF3)) <= 2e-9


# --------------------------------------------------------------------------
# Per-region tables (evaluation_optimized.c:159-188, :360-422)
# --------------------------------------------------------------------------

R_STEPS, R_MIN, R_MAX = 70, 16.0, 22.5
NU_STEPS, NU_MIN, NU_MAX = 64, -1.25, 1.25
MU_STEP = 100.0 / 160


@pytest.fixture(scope="module")
def tables():
    return O.r_constants(30, R_STEPS, R_MIN, (R_MAX - R_MIN) / R_STEPS, MU_STEP,
                         NU_STEPS, NU_MIN, (NU_MAX - NU_MIN) / NU_STEPS)


def _r_kpc(g):
    return 10 ** ((np.asarray(g) - 14.2) / 5)


def test_radial_volumes_telescope(tables):
    """sum_k irv_k = (r_max^3 - r_min^3) / 3 * dmu (radians), r = 10^((g - 14.2) / 5)."""
    want = (_r_kpc(R_MAX) ** 3 - _r_kpc(R_MIN) ** 3) / 3 * MU_STEP * math.pi / 180
    assert rel(tables["irv"].sum(), want) <= 1e-12
    assert np.all(tables["irv"] > 0)


def test_nu_slabs_telescope(tables):
    """sum_j ids_j = sin(nu_max) - sin(nu_min); nus are the slab mid-points."""
    want = math.sin(math.radians(NU_MAX)) - math.sin(math.radians(NU_MIN))
    assert rel(tables["ids"].sum(), want) <= 1e-12
    dnu = (NU_MAX - NU_MIN) / NU_STEPS
    assert np.allclose(tables["nus"], NU_MIN + (np.arange(NU_STEPS) + 0.5) * dnu, rtol=0, atol=1e-14)


def test_convolution_distances_closed_form(tables):
    """r_point_i = r' * 10^(dx_i / 5) at the r-step mid-point r', and
    r_in_mag = g' + dx_i, r_in_mag2 = its square."""
    g = R_MIN + np.arange(R_STEPS) * (R_MAX - R_MIN) / R_STEPS
    rp = (_r_kpc(g) + _r_kpc(g + (R_MAX - R_MIN) / R_STEPS)) / 2
    x, _ = leggauss(30)
    dx = 3 * 0.6 * x
    assert rel(tables["r_point"], rp[:, None] * 10 ** (dx[None, :] / 5)) <= 1e-13
    gp = 5 * np.log10(1000 * rp) - 5 + 4.2
    assert np.max(np.abs(tables["r_in_mag"] - (gp[:, None] + dx[None, :]))) <= 1e-12
    assert rel(tables["r_in_mag2"], tables["r_in_mag"] ** 2) <= 1e-15


def test_detection_efficiency_sigmoid(tables):
    """reff_xr_rp3 * r'^3 / (3 sigma) is the sigmoid 0.9402 / (1 + exp(1.6171 (g' - 23.5877)))."""
    g = R_MIN + np.arange(R_STEPS) * (R_MAX - R_MIN) / R_STEPS
    rp = (_r_kpc(g) + _r_kpc(g + (R_MAX - R_MIN) / R_STEPS)) / 2
    gp = 5 * np.log10(1000 * rp) - 1 * 5 + 4.2
    reff = 0.9402 / (1 + np.exp(1.6171 * (gp - 23.5877)))
    assert rel(tables["reff_xr_rp3"] * rp ** 3 / (3 * 0.6), reff) <= 1e-13
    # the detection efficiency falls through the faint end
    assert reff[0] > 0.9 and reff[-1] < reff[0]


# --------------------------------------------------------------------------
# Survey geometry (atSurveyGeometry.c, stCoords.c)
# --------------------------------------------------------------------------

def test_wedge_inclination_and_node():
    assert O.wedge_incl(86) == pytest.approx(10.0, abs=1e-13)   # 2.5*86 - 57.5 - 180 + 32.5
    assert O.wedge_incl(10) == pytest.approx(0.0, abs=1e-13)
    assert O.wedge_incl(46) == pytest.approx(90.0, abs=1e-13)   # 46 is not > 46
    assert O.get_node() == 95.0


def test_node_lies_on_the_equator():
    ra, dec = O.gc_to_eq(95.0, 0.0, 95.0, 37.0)
    assert ra == pytest.approx(95.0, abs=1e-12) and abs(dec) <= 1e-12


@pytest.mark.parametrize("wedge", [10, 79, 86])
def test_great_circle_plane(wedge):
    """Points with nu = 0 lie on the great circle whose pole is the node frame's
    z axis tilted by the wedge inclination."""
    inc = math.radians(O.wedge_incl(wedge))
    pole = PM.rot_z(math.radians(95.0)) @ PM.rot_x(inc) @ np.array([0, 0, 1.0])
    for mu in np.linspace(0, 360, 13):
        ra, dec = O.gc_to_eq(mu, 0.0, 95.0, O.wedge_incl(wedge))
        v = np.array([math.cos(math.radians(dec)) * math.cos(math.radians(ra)),
                      math.cos(math.radians(dec)) * math.sin(math.radians(ra)),
                      math.sin(math.radians(dec))])
        assert abs(np.dot(v, pole)) <= 1e-13


@pytest.mark.parametrize("wedge", [9, 12, 79, 86])
def test_sdss_chain_matches_rotation_matrices(wedge):
    mu = np.linspace(-40, 420, 47)
    for nu in (-1.25, -0.3, 0.0, 0.9):
        for m in mu:
            l, b = O.gc_to_lb(wedge, 0, m, nu)
            L, B = PM.gc_to_lb(wedge, 0, np.array(m), np.array(nu))
            assert abs((l - float(L) + 180) % 360 - 180) <= 1e-11
            assert abs(b - float(B)) <= 1e-11
            assert 0 <= l < 360 and -90 <= b <= 90


def test_galactic_matrix_is_orthonormal():
    M = PM.GAL_MATRIX
    assert np.max(np.abs(M @ M.T - np.eye(3))) <= 1e-10   # a 12-digit table


@pytest.mark.parametrize("wedge", [10, 15, 82])
def test_sgr_chain_matches_majewski_rotation(wedge):
    checked = 0
    for m in np.linspace(-40, 420, 47):
        for nu in (-1.2, 0.0, 0.7):
            lam, beta = O.gc_to_sgr(m, nu, wedge)
            if lam > 360:
                continue
            l, b = O.gc_to_lb(wedge, 1, m, nu)
            L, B = PM.gc_to_lb(wedge, 1, np.array(m), np.array(nu))
            assert abs((l - float(L) + 180) % 360 - 180) <= 1e-10
            assert abs(b - float(B)) <= 1e-10
            checked += 1
    assert checked > 50


def test_sgr_longitudes_above_360_quirk():
    """gcToSgr wraps only negative longitudes, and sgrToGal's quadrant cases
    stop at 360, so lambda > 360 leaves the in-plane components at zero
    (atSurveyGeometry.c:261-264, :295-386): the result is the image of the
    Sagittarius pole direction scaled by cos(beta + 90)."""
    hits = 0
    for m in np.linspace(95, 265, 18):
        lam, beta = O.gc_to_sgr(m, -1.2, 82)
        if lam <= 360:
            continue
        hits += 1
        z2 = math.cos(math.radians(beta + 90))
        g = PM.sgr_rotation().T[:, 2] * z2
        l, b = O.gc_to_lb(82, 1, m, -1.2)
        assert b == pytest.approx(math.degrees(math.asin(g[2])), abs=1e-10)
        assert l == pytest.approx(math.degrees(math.atan2(g[1], g[0])) % 360, abs=1e-10)
    assert hits > 5


def test_lbr2xyz_cancelling_form():
    for l, b, r in [(0.0, 0.0, 8.5), (30.0, 45.0, 12.0), (200.0, -60.0, 40.0), (359.0, 89.0, 1.0)]:
        x, y, z = O.lbr2xyz(l, b, r)
        lr, br = math.radians(l), math.radians(b)
        assert x == pytest.approx(r * math.cos(br) * math.cos(lr) - 8.5, abs=1e-12)
        assert y == pytest.approx(r * math.cos(br) * math.sin(lr), abs=1e-13)
        assert z == pytest.approx(r * math.sin(br), abs=1e-13)
    assert np.allclose(O.lbr2xyz(123.0, 45.0, 0.0), [-8.5, 0.0, 0.0], atol=1e-14)


# --------------------------------------------------------------------------
# Integrand special cases (evaluation_optimized.c:190-299)
# --------------------------------------------------------------------------

def _one_stream_model(theta, phi, sigma=2.0, mu=350.0, r=20.0, q=0.7, r0=6.0):
    # n = 10: the reference's quadrature weights are exact to ~3e-12 there
    p, _ = synth.workload(30, synth.WedgeShape(5, 5, 5, 10, 10), 1)
    p.background_parameters[1] = q
    p.background_parameters[2] = r0
    p.streams[0].parameters = [mu, r, theta, phi, sigma]
    return p


def test_stream_axis_along_line_of_sight():
    """A stream whose centre lies on a star's line of sight and whose axis is
    that line: every convolution point is on the axis, so the stream density
    is exp(0) = 1 and st_prob = reff_xr_rp3 * sum_i qw_i."""
    mu, r_s = 350.0, 20.0
    l, b = O.gc_to_lb(86, 0, mu, 0.0)
    u = PM.unit_los(l, b)
    theta, phi = math.acos(u[2]), math.atan2(u[1], u[0])
    p = _one_stream_model(theta, phi, mu=mu, r=r_s)
    ph = ParamsHolder(p)
    r_star = 15.0
    _, st = O.point_probabilities(ph, l, b, r_star)
    rho, ri, qw, _ = PM.distance_tables(np.array(r_star), p.convolve)
    assert rel(st[0], rho * np.sum(qw)) <= 1e-11


def test_disabled_stream_and_zero_q():
    p = _one_stream_model(1.0, 2.0, sigma=5e-5)
    bg, st = O.point_probabilities(ParamsHolder(p), 200.0, 50.0, 10.0)
    assert st[0] == 0.0 and bg > 0
    p = _one_stream_model(1.0, 2.0, q=0.0)
    bg, st = O.point_probabilities(ParamsHolder(p), 200.0, 50.0, 10.0)
    assert bg == -1.0 and st[0] > 0


def test_background_only_likelihood_identity():
    """No streams: P = 1000 * B / I_b, so log10 P = 3 + log10(B / I_b)."""
    p, stars = synth.workload(31, synth.WedgeShape(8, 6, 4, 12, 40), 0)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    res = O.evaluate_full(ph, sh)
    for i, (l, b, r) in enumerate(stars):
        B, _ = O.point_probabilities(ph, l, b, r)
        assert res.log10_p[i] == pytest.approx(3 + math.log10(B / res.background_integral), abs=1e-13)


def test_zero_probability_stars_count_minus_238():
    """exp(-800) == 0 background weight and a narrow stream far from the stars:
    P == 0 exactly, which the reference counts as -238 and in num_zero."""
    p, stars = synth.workload(32, synth.WedgeShape(8, 6, 4, 12, 30), 1)
    p.background_weight = -800.0
    p.streams[0].parameters = [350.0, 40.0, 0.3, 1.0, 0.05]
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    res = O.evaluate_full(ph, sh)
    zero = res.log10_p == -238.0
    assert zero.sum() == res.num_zero and res.num_zero > 0
    assert res.prob_sum == pytest.approx(np.sum(res.log10_p), abs=1e-9)


# --------------------------------------------------------------------------
# Whole evaluation against the independent model
# --------------------------------------------------------------------------

def _check_model(p, stars, tol=2e-12):
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    m = PM.evaluate(p, stars)
    assert rel(ref.background_integral, m["I_bg"]) <= tol
    if p.number_streams:
        assert rel(ref.stream_integrals, m["I_st"]) <= tol
    for (bg, st), (mbg, mst) in zip(ref.regions, m["regions"]):
        assert rel(bg, mbg) <= tol
        if p.number_streams:
            assert rel(st, mst) <= tol
    assert np.max(np.abs(ref.log10_p - m["log10p"])) <= tol
    assert abs(ref.likelihood - m["likelihood"]) <= tol


@pytest.mark.parametrize("S", [0, 1, 2, 4])
def test_model_synthetic(S):
    p, stars = synth.workload(40 + S, synth.WedgeShape(12, 10, 6, 9, 300), S)
    _check_model(p, stars)


def test_model_general_alpha_delta():
    p, stars = synth.workload(45, synth.WedgeShape(12, 10, 6, 9, 300), 2)
    p.background_parameters[0] = 1.4
    p.background_parameters[3] = 0.7
    _check_model(p, stars)


@pytest.mark.parametrize("fid,aux", [("79", False), ("86", False), ("11", False), ("86", True)])
def test_model_reference_fixtures(fid, aux):
    p, stars = fixture(fid, n_stars=300, aux=aux, convolve=10)
    if fid == "11":
        p.regions[0].mu_steps, p.regions[0].mu_max = 40, p.regions[0].mu_min + 40 * (
            p.regions[0].mu_max - p.regions[0].mu_min) / 160
    _check_model(p, stars)


def test_model_sagittarius_wedge():
    p, stars = synth.workload(46, synth.WedgeShape(12, 10, 6, 9, 300), 1)
    p.sgr_coordinates, p.wedge = 1, 10
    for g in p.regions:
        g.mu_min, g.mu_max = -20.0, 30.0
    p.streams[0].parameters[0] = 5.0
    _check_model(p, stars)


# --------------------------------------------------------------------------
# Evaluator plumbing
# --------------------------------------------------------------------------

def test_openmp_baseline_matches_serial():
    p, stars = synth.workload(47, synth.WedgeShape(20, 24, 8, 15, 500), 2)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    a = O.evaluate_full(ph, sh)
    b = O.evaluate_full(ph, sh, threads=4)
    assert rel(a.background_integral, b.background_integral) <= 1e-13
    assert rel(a.stream_integrals, b.stream_integrals) <= 1e-13
    assert abs(a.likelihood - b.likelihood) <= 1e-13


@pytest.mark.parametrize("fid", ["11", "79", "86"])
def test_parameter_vector_mapping(fid):
    p, _ = fixture(fid, n_stars=10)
    vec = ctest_vectors()[fid]
    assert p.optimized_parameter_count() == len(vec)
    assert np.array_equal(p.parameter_vector(), np.asarray(vec))
    ph = ParamsHolder(p)
    assert O.lib().orc_get_optimized_parameter_count(ph.ptr) == len(vec)


def test_evaluate_return_value_and_state():
    p, stars = fixture("79", n_stars=200, convolve=10)
    vec = p.parameter_vector()
    ph, sh, st = holders(p, stars)
    v = O.lib().orc_evaluate(np.ascontiguousarray(vec).ctypes.data_as(O._dp), ph.ptr, st.ptr, sh.ptr)
    assert v == pytest.approx(st.es.prob_sum / len(stars), abs=0)
    assert st.es.current_star_point == len(stars)
    assert st.area(0).current_calculation == st.area(0).max_calculation


def test_split_range_equals_one_shot():
    """min/max_calculation range split (the reference's MPI split) sums to the
    one-shot integral."""
    p, stars = synth.workload(48, synth.WedgeShape(10, 12, 6, 8, 50), 1)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    whole = O.integrals(ph, sh)
    st = StateHolder(ph, sh)
    total = st.area(0).max_calculation
    st.area(0).max_calculation = 301
    O.integrals(ph, sh, st)
    st.es.current_integral = 0
    st.area(0).max_calculation = total
    O.integrals(ph, sh, st)
    assert rel(st.es.background_integral, whole.es.background_integral) <= 1e-14
    assert rel(st.stream_integrals(), whole.stream_integrals()) <= 1e-14
