"""Mixed precision against the FP64 oracle.

The mixed path evaluates the integrand's sqrt, reciprocal and exp in FP32
and keeps FP64 for positions, the stream-axis distance, all sums and the
per-star log.  Bar: every star's |delta ln P| <= 1e-7 against the FP64
oracle; the integrals' relative errors are reported and held to 1e-8.
"""
import math

import numpy as np
import pytest

from helpers import fixture, rel
from milkyway_client_amd import separation as sep
from milkyway_client_amd import synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LNP_TOL = 1e-7        # per star, natural log
INT_REL_TOL = 1e-8


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


def mixed_errors(p, stars, vectors=None):
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    vecs = [p.parameter_vector()] if vectors is None else list(vectors)
    out = []
    with sep.Context(ph, sh) as ctx:
        got = ctx.eval(np.array(vecs), precision=sep.PREC_MIXED, per_star=True)
    for v, g in zip(vecs, got):
        p.set_parameter_vector(v)
        ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
        d_lnp = math.log(10) * np.max(np.abs(g.log10_p - ref.log10_p))
        out.append(dict(
            I_bg=rel(g.background_integral, ref.background_integral),
            I_st=rel(g.stream_integrals, ref.stream_integrals) if p.number_streams else 0.0,
            lnP=d_lnp, lnL=abs(g.likelihood - ref.likelihood) * math.log(10)))
    return out


@pytest.mark.parametrize("fid", ["79", "86", "11"])
def test_mixed_reference_fixtures(fid):
    p, stars = fixture(fid, n_stars=2000)
    for e in mixed_errors(p, stars):
        print(fid, {k: f"{v:.2e}" for k, v in e.items()})
        assert e["lnP"] <= LNP_TOL
        assert e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL


def test_mixed_sweep_parameter_vectors():
    """32 seeded parameter vectors on a coarser grid over the full-size
    wedge's g, mu and nu spans (70 x 80 x 32 cells, n = 120, 3 streams)."""
    shape = synth.WedgeShape(70, 80, 32, 120, 3000)
    p, stars = synth.workload(1, shape, 3)
    vecs = synth.parameter_vectors(2, p, 32)
    errs = mixed_errors(p, stars, vecs)
    worst = {k: max(e[k] for e in errs) for k in errs[0]}
    print("worst over 32 vectors:", {k: f"{v:.2e}" for k, v in worst.items()})
    assert worst["lnP"] <= LNP_TOL
    assert worst["I_bg"] <= INT_REL_TOL and worst["I_st"] <= INT_REL_TOL


@pytest.mark.parametrize("n_conv", [7, 9, 17, 64, 100, 101, 107, 120])
def test_mixed_group_remainders(n_conv):
    """Rows of at least 100 convolution points sum their stream terms in
    FP32 groups of 8 points, the points left at a row's end in a shorter
    group; coarser rows add each term in FP64.  Counts on both sides of the
    switch and with 4, 5, 3 and 0 points left over, 4 streams, 4 seeded
    vectors each, against the oracle."""
    shape = synth.WedgeShape(40, 48, 24, n_conv, 1500)
    p, stars = synth.workload(5, shape, 4)
    errs = mixed_errors(p, stars, synth.parameter_vectors(7, p, 4))
    worst = {k: max(e[k] for e in errs) for k in errs[0]}
    print(f"n = {n_conv}:", {k: f"{v:.2e}" for k, v in worst.items()})
    assert worst["lnP"] <= LNP_TOL
    assert worst["I_bg"] <= INT_REL_TOL and worst["I_st"] <= INT_REL_TOL


@pytest.mark.parametrize("n_conv, streams", [(128, 0), (128, 1), (200, 2), (150, 3)])
def test_mixed_group_sums_other_stream_counts(n_conv, streams):
    """The grouped rows with no stream (the background alone) and with 1, 2
    and 3 streams, at convolutions beyond the reference files' 120 points
    (200 fills the LDS tile with five rows), 4 seeded vectors each, against
    the oracle."""
    shape = synth.WedgeShape(40, 48, 24, n_conv, 1500)
    p, stars = synth.workload(11 + streams, shape, streams)
    errs = mixed_errors(p, stars, synth.parameter_vectors(13, p, 4))
    worst = {k: max(e[k] for e in errs) for k in errs[0]}
    print(f"n = {n_conv}, S = {streams}:", {k: f"{v:.2e}" for k, v in worst.items()})
    assert worst["lnP"] <= LNP_TOL
    assert worst["I_bg"] <= INT_REL_TOL and worst["I_st"] <= INT_REL_TOL


def _aux_79():
    return fixture("79", n_stars=1000, aux=True, convolve=30)


def _general_alpha_delta():
    p, stars = synth.workload(3, synth.WedgeShape(20, 24, 16, 15, 500), 2)
    p.background_parameters[0] = 1.3
    p.background_parameters[3] = 0.8
    return p, stars


def _zero_q():
    p, stars = synth.workload(3, synth.WedgeShape(20, 24, 16, 15, 500), 2)
    p.background_parameters[1] = 0.0
    return p, stars


@pytest.mark.parametrize("model", [_aux_79, _general_alpha_delta, _zero_q],
                         ids=["aux_profile", "general_alpha_delta", "zero_q"])
def test_mixed_request_on_fp64_backgrounds(model):
    """The mixed kernels cover the Hernquist background with alpha = delta =
    1 (the reference files' and the benchmark's); for the aux profile,
    general alpha and delta, and q = 0 both passes run the FP64 kernels.  A
    mixed request on those models returns the FP64 result bitwise, and that
    is the oracle's at the FP64 bars (integrals 1e-12, ln L and every star
    1e-10)."""
    p, stars = model()
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    vec = p.parameter_vector()
    with sep.Context(ph, sh) as ctx:
        f = ctx.eval(vec, precision=sep.PREC_FP64, per_star=True)[0]
        m = ctx.eval(vec, precision=sep.PREC_MIXED, per_star=True)[0]
    assert (m.likelihood, m.background_integral, m.num_zero) == (f.likelihood, f.background_integral, f.num_zero)
    assert np.array_equal(m.stream_integrals, f.stream_integrals)
    assert np.array_equal(m.log10_p, f.log10_p)
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    print(model.__name__, rel(f.background_integral, ref.background_integral),
          rel(f.stream_integrals, ref.stream_integrals), abs(f.likelihood - ref.likelihood))
    assert rel(f.background_integral, ref.background_integral) <= 1e-12
    assert rel(f.stream_integrals, ref.stream_integrals) <= 1e-12
    assert abs(f.likelihood - ref.likelihood) <= 1e-10
    assert np.max(np.abs(f.log10_p - ref.log10_p)) <= 1e-10
