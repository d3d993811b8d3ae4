"""Models of more than 4 streams against the oracle.

The reference sizes everything by ap->number_streams
(evaluation_optimized.c:56-60, 87, 276-298, 639-701): calculate_integrals and
calculate_likelihood take any count.  The library evaluates such a model in
passes of 4 streams (the background in the first), chaining each star's
mixture sum from pass to pass in stream order.  Bars as for the FP64 parity
tests: integrals 1e-12 relative, the likelihood and every star's log10 P
1e-10; mixed precision per star 1e-7 in ln P.
"""
import math

import numpy as np
import pytest

from helpers import holders, rel
from milkyway_client_amd import separation as sep
from milkyway_client_amd import synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder
from oracle import oracle as O
from test_gpu_parity import INT_TOL, LNL_TOL, check_against_oracle

pytestmark = pytest.mark.gpu

CUT_SHAPE = synth.WedgeShape(25, 36, 14, 20, 600, cut=(25, 10, 14, 330.0, 345.0))


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


@pytest.mark.parametrize("S", [5, 6, 8, 9])
def test_context_many_streams(S):
    p, stars = synth.workload(200 + S, CUT_SHAPE, S)
    res, ref = check_against_oracle(p, stars)
    assert res.stream_integrals.size == S


def test_dropin_evaluate_six_streams():
    p, stars = synth.workload(206, CUT_SHAPE, 6)
    vec = p.parameter_vector()
    ph, sh, es = holders(p, stars)
    got = sep.evaluate(ph, es, sh, vec)
    ph2, sh2, _ = holders(p, stars)
    want = O.evaluate(ph2, sh2, vec)
    assert abs(got - want) <= LNL_TOL
    ph3, sh3, es3 = holders(p, stars)
    O.integrals(ph3, sh3, es3)
    assert rel(es.es.background_integral, es3.es.background_integral) <= INT_TOL
    assert rel(es.stream_integrals(), es3.stream_integrals()) <= INT_TOL
    for g in range(2):
        bg, st = es.region_integrals(g)
        bg3, st3 = es3.region_integrals(g)
        assert rel(bg, bg3) <= INT_TOL and rel(st, st3) <= INT_TOL
    assert es.es.current_star_point == len(stars)


def test_dropin_split_ranges_five_streams():
    """Partial integral ranges and a resumed star pass with 5 streams."""
    p, stars = synth.workload(205, synth.WedgeShape(20, 30, 12, 15, 500), 5)
    ph, sh, es = holders(p, stars)
    ph2, sh2, es2 = holders(p, stars)
    total = es.area(0).max_calculation
    k = total // 3 + 5
    es.area(0).max_calculation = k
    sep.calculate_integrals(ph, es, sh)
    es.es.current_integral = 0
    es.area(0).max_calculation = total
    sep.calculate_integrals(ph, es, sh)
    O.integrals(ph2, sh2, es2)
    assert rel(es.es.background_integral, es2.es.background_integral) <= INT_TOL
    assert rel(es.stream_integrals(), es2.stream_integrals()) <= INT_TOL
    es.es.current_star_point = 123
    es2.es.current_star_point = 123
    sep.calculate_likelihood(ph, es, sh)
    O.likelihood(ph2, sh2, es2)
    assert abs(es.es.prob_sum - es2.es.prob_sum) <= LNL_TOL * len(stars)


def test_mixed_precision_six_streams():
    p, stars = synth.workload(216, CUT_SHAPE, 6)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    ref = O.evaluate_full(ph, sh)
    with sep.Context(ph, sh) as ctx:
        got = ctx.eval(p.parameter_vector(), precision=sep.PREC_MIXED, per_star=True)[0]
        rs = ctx.eval(p.parameter_vector(), precision=sep.PREC_RESAMPLED, per_star=True)[0]
    for r in (got, rs):     # more than 4 streams run on the exact quadrature in both modes
        assert math.log(10) * np.max(np.abs(r.log10_p - ref.log10_p)) <= 1e-7
        assert rel(r.stream_integrals, ref.stream_integrals) <= 1e-7
    assert rs.likelihood == got.likelihood


def test_batch_and_server_refuse_more_than_four_streams():
    p, stars = synth.workload(207, synth.WedgeShape(10, 12, 8, 10, 50), 5)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        with pytest.raises(sep.MwError):
            sep.Batch([(ctx, 1)])
        with pytest.raises(sep.MwError):
            ctx.separation()
    with sep.Server(max_inflight=1, max_batch=1) as srv:
        with pytest.raises(sep.MwError):
            srv.add_wedge(ph, sh)
