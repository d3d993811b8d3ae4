"""Corners of the wedge and of the model, against the oracle in every
precision (evaluation_optimized.c:190-299, :424-753).

Each case changes one thing the synthetic workloads never reach: a wedge of a
single sight, a wedge of one r-step, a convolution of one point and one of
1024 (the most the library takes: one r-step per LDS tile), a stream whose axis runs along the
wedge's central sight (the sights near it are almost parallel to the axis),
a narrow stream centred 0.7 kpc from the Sun, and a model at the far edges
of the reference's parameter bounds together with a second stream that
passes outside the wedge.  Bars: FP64 integrals within 1e-12 relative and
every star's ln P within 1e-10; mixed and resampled integrals within 2e-8
and every star's ln P within 1e-7.
"""
import math

import numpy as np
import pytest

from helpers import rel
from milkyway_client_amd import separation as sep
from milkyway_client_amd import synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BARS = {sep.PREC_FP64: (1e-12, 1e-10), sep.PREC_MIXED: (2e-8, 1e-7), sep.PREC_RESAMPLED: (2e-8, 1e-7)}
LONG_R = synth.WedgeShape(700, 24, 16, 120, 2000)


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


def axis_along_central_sight(p):
    """Stream 0's axis along the unit vector of the sight at the wedge's
    centre (mu at mid-span, nu = 0), its centre 20 kpc out on that sight."""
    g0 = p.regions[0]
    mu_c = 0.5 * (g0.mu_min + g0.mu_max)
    l, b = synth.gc_to_lb(synth.WEDGE, mu_c, 0.0)
    l, b = math.radians(float(l)), math.radians(float(b))
    u = (math.cos(b) * math.cos(l), math.cos(b) * math.sin(l), math.sin(b))
    st = p.streams[0].parameters
    st[0], st[1], st[4] = mu_c, 20.0, 1.0
    st[2] = math.acos(u[2])
    st[3] = math.atan2(u[1], u[0])


def narrow_at_the_sun(p):
    st = p.streams[0].parameters
    st[1], st[4] = 0.7, 0.1


def bound_edges(p):
    """q and r0 at their largest (1.0 and 30 kpc), stream 0 at sigma 20 kpc
    and its largest centre distance (57.5 kpc), stream 1 passing outside the
    wedge (centre 140 kpc out)."""
    p.background_parameters[1] = 1.0
    p.background_parameters[2] = 30.0
    s0, s1 = p.streams[0].parameters, p.streams[1].parameters
    s0[1], s0[4] = 57.5, 20.0
    s1[1], s1[2], s1[3], s1[4] = 140.0, 1.047, 1.5708, 1.0


CASES = {
    "single_sight": (synth.WedgeShape(40, 1, 1, 30, 300), 2, None),
    "one_convolution_point": (synth.WedgeShape(40, 8, 4, 1, 300), 2, None),
    "axis_along_the_central_sight": (LONG_R, 1, axis_along_central_sight),
    "narrow_stream_at_the_sun": (LONG_R, 2, narrow_at_the_sun),
    "bound_edges_and_a_far_stream": (LONG_R, 2, bound_edges),
    "single_r_step": (synth.WedgeShape(1, 30, 12, 30, 300), 2, None),
    "convolution_at_its_maximum": (synth.WedgeShape(6, 6, 4, 1024, 300), 2, None),
}


@pytest.mark.parametrize("case", list(CASES))
def test_corner_against_the_oracle(case):
    shape, S, change = CASES[case]
    p, stars = synth.workload(401, shape, S)
    if change:
        change(p)
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    assert np.all(np.isfinite(ref.log10_p))
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        for prec, (int_tol, lnp_tol) in BARS.items():
            got = ctx.eval(p.parameter_vector(), precision=prec, per_star=True)[0]
            e = dict(I_bg=rel(got.background_integral, ref.background_integral),
                     I_st=rel(got.stream_integrals, ref.stream_integrals),
                     lnP=math.log(10) * float(np.max(np.abs(got.log10_p - ref.log10_p))),
                     lnL=math.log(10) * abs(got.likelihood - ref.likelihood))
            print(case, prec, {k: f"{v:.2e}" for k, v in e.items()})
            assert e["I_bg"] <= int_tol and e["I_st"] <= int_tol, (case, prec, e)
            assert e["lnP"] <= lnp_tol and e["lnL"] <= lnp_tol, (case, prec, e)


def test_convolution_beyond_its_maximum_is_refused():
    """1025 convolution points exceed the LDS r-tile: the context refuses
    the model with MW_E_INVALID and says why."""
    p, stars = synth.workload(401, synth.WedgeShape(4, 4, 4, 1025, 50), 1)
    with pytest.raises(sep.MwError) as e:
        sep.Context(ParamsHolder(p), StarsHolder(stars))
    assert e.value.code == sep.MW_E_INVALID
    assert "convolve" in str(e.value)
