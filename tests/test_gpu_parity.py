"""Parity of the HIP library (through its C ABI) with the CPU oracle.

Bars (FP64 path): every integral within 1e-12 relative, |delta lnL| <= 1e-10
on the evaluator's result (mean log10 P), and every star's log10 P within
1e-10 absolute.  Inputs: the reference's own test fixtures (parameter files,
star subsets, CTest parameter vectors) and seeded synthetic wedges.
"""
import math

import numpy as np
import pytest

from helpers import fixture, holders, rel
from milkyway_client_amd import separation as sep
from milkyway_client_amd import synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder
from oracle import oracle as O

pytestmark = pytest.mark.gpu

INT_TOL = 1e-12
LNL_TOL = 1e-10
STAR_TOL = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


def check_against_oracle(p, stars, precision=sep.PREC_FP64, int_tol=INT_TOL, lnl_tol=LNL_TOL,
                         star_tol=STAR_TOL):
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    ref = O.evaluate_full(ph, sh)
    with sep.Context(ph, sh) as ctx:
        vec = p.parameter_vector()
        res = ctx.eval(vec, precision=precision, per_star=True, regions=True)[0]
    S = p.number_streams
    print(f"I_bg gpu={res.background_integral!r} oracle={ref.background_integral!r}")
    assert rel(res.background_integral, ref.background_integral) <= int_tol
    assert rel(res.stream_integrals, ref.stream_integrals) <= int_tol
    for g, (bg, st) in enumerate(ref.regions):
        assert rel(res.regions[g][0], bg) <= int_tol
        assert rel(res.regions[g][1:1 + S], st) <= int_tol
    print(f"lnL gpu={res.likelihood!r} oracle={ref.likelihood!r}")
    assert abs(res.likelihood - ref.likelihood) <= lnl_tol
    assert res.num_zero == ref.num_zero
    assert np.max(np.abs(res.log10_p - ref.log10_p)) <= star_tol
    assert abs(res.bg_only_likelihood - ref.bg_only_mean) <= lnl_tol
    if S:
        assert np.max(np.abs(res.stream_only_likelihood - ref.st_only_mean)) <= lnl_tol
    return res, ref


def test_reference_size_synthetic():
    p, stars = synth.workload(0, synth.REFERENCE_SIZE, 1)
    check_against_oracle(p, stars)


@pytest.mark.parametrize("fid", ["79", "86", "11"])
def test_reference_fixture(fid):
    p, stars = fixture(fid, n_stars=2000)
    check_against_oracle(p, stars)


def test_fixture_12_convolve_30():
    p, stars = fixture("12", n_stars=1000, convolve=30)
    check_against_oracle(p, stars)


@pytest.mark.parametrize("fid", ["79", "86"])
def test_aux_profile_fixture(fid):
    p, stars = fixture(fid, n_stars=1000, aux=True, convolve=30)
    assert p.aux_bg_profile == 1
    check_against_oracle(p, stars)


def test_general_alpha_delta():
    p, stars = synth.workload(3, synth.WedgeShape(20, 24, 16, 15, 500), 2)
    p.background_parameters[0] = 1.3
    p.background_parameters[3] = 0.8
    check_against_oracle(p, stars)


@pytest.mark.parametrize("S", [0, 2, 4])
def test_stream_counts(S):
    p, stars = synth.workload(10 + S, synth.WedgeShape(30, 40, 20, 20, 700), S)
    check_against_oracle(p, stars)


def test_disabled_stream_sigma():
    # |sigma| < 1e-4 switches a stream off (evaluation_optimized.c:279-280);
    # the reference then divides 0 by a zero integral for the star pass, so
    # only the integrals are compared.
    p, stars = synth.workload(21, synth.WedgeShape(20, 30, 10, 15, 300), 2)
    p.streams[1].parameters[4] = 5e-5
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    st = O.integrals(ph, sh)
    with sep.Context(ph, sh) as ctx:
        res = ctx.eval(p.parameter_vector())[0]
    assert res.stream_integrals[1] == 0.0
    assert rel(res.background_integral, st.es.background_integral) <= INT_TOL
    assert rel(res.stream_integrals[0], st.stream_integrals()[0]) <= INT_TOL


def test_dropin_evaluate_matches_oracle():
    p, stars = fixture("79", n_stars=2000)
    vec = p.parameter_vector()
    ph, sh, es = holders(p, stars)
    got = sep.evaluate(ph, es, sh, vec)
    ph2, sh2, _ = holders(p, stars)
    want = O.evaluate(ph2, sh2, vec)
    assert abs(got - want) <= LNL_TOL
    # side effects on the evaluation state
    assert es.es.current_integral == ph.ap.number_integrals
    assert es.es.current_star_point == len(stars)
    a = es.area(0)
    assert a.current_calculation == a.max_calculation


def test_dropin_split_ranges_and_resume():
    """The integral honours current/max_calculation: two partial calls over
    [0, k) and [k, K) give the one-shot result, as the MPI range split and
    checkpoint resume of the reference rely on."""
    p, stars = synth.workload(5, synth.WedgeShape(25, 30, 12, 20, 200), 1)
    ph, sh, es = holders(p, stars)
    ref_state = O.integrals(ph, sh)
    total = es.area(0).max_calculation
    k = total // 3 + 7            # not on a column boundary
    es.area(0).max_calculation = k
    sep.calculate_integrals(ph, es, sh)
    assert es.area(0).current_calculation == k
    es.es.current_integral = 0
    es.area(0).max_calculation = total
    sep.calculate_integrals(ph, es, sh)
    assert es.area(0).current_calculation == total
    assert rel(es.es.background_integral, ref_state.es.background_integral) <= INT_TOL
    assert rel(es.stream_integrals(), ref_state.stream_integrals()) <= INT_TOL


def test_dropin_partial_range_matches_oracle_partial():
    p, stars = synth.workload(6, synth.WedgeShape(15, 20, 10, 10, 100), 2)
    ph, sh, es = holders(p, stars)
    ph2, sh2, es2 = holders(p, stars)
    for st in (es, es2):
        st.area(0).current_calculation = 123
        st.area(0).max_calculation = 2345
    sep.calculate_integrals(ph, es, sh)
    O.integrals(ph2, sh2, es2)
    assert rel(es.es.background_integral, es2.es.background_integral) <= INT_TOL
    assert rel(es.stream_integrals(), es2.stream_integrals()) <= INT_TOL
    assert es.area(0).current_calculation == es2.area(0).current_calculation


def test_likelihood_resume_from_star():
    p, stars = fixture("79", n_stars=600)
    ph, sh, es = holders(p, stars)
    ph2, sh2, es2 = holders(p, stars)
    sep.calculate_integrals(ph, es, sh)
    O.integrals(ph2, sh2, es2)
    es.es.current_star_point = 250
    es2.es.current_star_point = 250
    sep.calculate_likelihood(ph, es, sh)
    O.likelihood(ph2, sh2, es2)
    assert abs(es.es.prob_sum - es2.es.prob_sum) <= LNL_TOL * 600
    assert es.es.current_star_point == 600


def test_batch_equals_single():
    """A batch of vectors gives each vector's single-evaluation result, bitwise."""
    p, stars = synth.workload(7, synth.WedgeShape(30, 40, 16, 30, 1000), 2)
    vecs = synth.parameter_vectors(7, p, 5)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        batch = ctx.eval(vecs)
        for v, b in zip(vecs, batch):
            one = ctx.eval(v)[0]
            assert one.likelihood == b.likelihood
            assert one.background_integral == b.background_integral
            assert np.array_equal(one.stream_integrals, b.stream_integrals)


def test_repeatable_bitwise():
    p, stars = synth.workload(8, synth.WedgeShape(30, 40, 16, 30, 1000), 1)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        a = ctx.eval(p.parameter_vector(), per_star=True)[0]
        b = ctx.eval(p.parameter_vector(), per_star=True)[0]
    assert a.likelihood == b.likelihood
    assert np.array_equal(a.log10_p, b.log10_p)


def test_sagittarius_wedges():
    """sgr_coordinates = 1, with mu ranges that reach Sagittarius longitudes
    above 360 (the reference's sgrToGal quirk, reproduced by both sides)."""
    for wedge, mu_lo, mu_hi in [(10, -20.0, 30.0), (82, 80.0, 200.0)]:
        p, stars = synth.workload(50 + wedge, synth.WedgeShape(20, 30, 12, 20, 400), 2)
        p.sgr_coordinates, p.wedge = 1, wedge
        for g in p.regions:
            g.mu_min, g.mu_max = mu_lo, mu_hi
        p.streams[0].parameters[0] = 0.5 * (mu_lo + mu_hi)
        check_against_oracle(p, stars)


def test_zero_probability_stars():
    """P == 0 stars count -238 and num_zero (evaluation_optimized.c:705-716)."""
    p, stars = synth.workload(60, synth.WedgeShape(20, 24, 8, 15, 400), 1)
    p.background_weight = -800.0
    p.streams[0].parameters = [350.0, 40.0, 0.3, 1.0, 0.05]
    res, ref = check_against_oracle(p, stars)
    assert ref.num_zero > 0


def test_full_size_geometry_slice():
    """The full-size wedge (700 x 1600 x 640 plus a 700 x 400 x 160 cut,
    n = 120, 3 streams): the same flat-index slice of every region through
    the drop-in entry point and through the oracle."""
    p, stars = synth.workload(1, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS, n_stars=2000)
    ph, sh, es = holders(p, stars)
    ph2, sh2, es2 = holders(p, stars)
    for st in (es, es2):
        a0 = st.area(0)
        a0.current_calculation = 700 * 640 * 811 + 12345   # mid-wedge, mid-column
        a0.max_calculation = a0.current_calculation + 2 * 640 * 700
        a1 = st.area(1)
        a1.current_calculation = 700 * 160 * 200
        a1.max_calculation = a1.current_calculation + 160 * 700
    sep.calculate_integrals(ph, es, sh)
    O.integrals(ph2, sh2, es2)
    for g in range(2):
        bg, st = es.region_integrals(g)
        bg2, st2 = es2.region_integrals(g)
        assert rel(bg, bg2) <= INT_TOL
        assert rel(st, st2) <= INT_TOL
    sep.calculate_likelihood(ph, es, sh)
    O.likelihood(ph2, sh2, es2)
    assert abs(es.es.prob_sum - es2.es.prob_sum) / len(stars) <= LNL_TOL


def test_zero_stars():
    """A star list of no stars.  The drop-in's integrals are the oracle's;
    the star pass runs no star and leaves prob_sum 0, num_zero 0 and
    current_star_point 0, and the evaluation returns prob_sum /
    (number_stars - bad_jacobians) = 0 / 0 = NaN, as cpu_evaluate does
    (evaluation_optimized.c:664-717, :756-784).  A context over the same
    list gives the integrals and a NaN likelihood."""
    p, stars = synth.workload(5, synth.WedgeShape(25, 30, 12, 20, 200), 1)
    none = np.zeros((0, 3))
    vec = p.parameter_vector()
    ph, sh, es = holders(p, none)
    got = sep.evaluate(ph, es, sh, vec)
    assert math.isnan(got)
    assert es.es.prob_sum == 0.0 and es.es.num_zero == 0 and es.es.current_star_point == 0
    ref = O.integrals(*holders(p, stars)[:2])
    assert rel(es.es.background_integral, ref.es.background_integral) <= INT_TOL
    assert rel(es.stream_integrals(), ref.stream_integrals()) <= INT_TOL
    with sep.Context(ParamsHolder(p), StarsHolder(none)) as ctx:
        for prec in (sep.PREC_FP64, sep.PREC_MIXED):
            r = ctx.eval(vec, precision=prec)[0]
            assert math.isnan(r.likelihood)
            assert rel(r.background_integral, ref.es.background_integral) <= (INT_TOL if prec == sep.PREC_FP64 else 1e-8)
