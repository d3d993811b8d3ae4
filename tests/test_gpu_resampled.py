"""The resampled-quadrature mode (MW_PREC_RESAMPLED) on the GPU.

It is mixed precision on a quadrature whose weights are folded onto a grid
in distance or magnitude.  Bars, as for the mixed path: every star's |delta ln P| <= 1e-7
against the FP64 oracle, integrals within 2e-8 relative (the FP32
transcendentals dominate; the quadrature change itself is ~1e-10).  Cases
include the narrowest streams the parameter ranges allow (sigma = 0.1 kpc),
the corners of the reference's parameter bounds against the oracle, every
stream count up to four, the aux profile, the full-size wedge against the
GPU's own FP64 result, and bitwise equality between single, batched
(eager and graph) and served evaluations.
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

LNP_TOL = 1e-7
INT_REL_TOL = 2e-8
LONG_R = synth.WedgeShape(700, 24, 16, 120, 2000)   # full r and convolution resolution


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


def errors_vs_oracle(p, stars):
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        before = ctx.resident_bytes
        got = ctx.eval(p.parameter_vector(), precision=sep.PREC_RESAMPLED, per_star=True)[0]
        grew = ctx.resident_bytes > before
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    e = dict(I_bg=rel(got.background_integral, ref.background_integral),
             I_st=rel(got.stream_integrals, ref.stream_integrals) if p.number_streams else 0.0,
             lnP=math.log(10) * np.max(np.abs(got.log10_p - ref.log10_p)),
             lnL=math.log(10) * abs(got.likelihood - ref.likelihood))
    print({k: f"{v:.2e}" for k, v in e.items()}, "resampled tables built:", grew)
    return e, grew


@pytest.mark.parametrize("S", [0, 1, 2, 3, 4])
def test_resampled_against_oracle(S):
    p, stars = synth.workload(90 + S, LONG_R, S)
    e, grew = errors_vs_oracle(p, stars)
    assert grew, "the resampled grid was not used"
    assert e["lnP"] <= LNP_TOL and e["lnL"] <= LNP_TOL
    assert e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL


def test_disabled_stream_under_the_recurrence():
    """A stream switched off by |sigma| < 1e-4 (evaluation_optimized.c:279-280)
    carries k_s = 0 in the model constants, so its recurrence is the
    constant 1 (no overflow) and the off flag zeroes it: the recurrence stays
    on, the disabled stream's integral is exactly 0 and the others match the
    oracle.  (The reference's star pass then divides 0 by the zero integral,
    so only the integrals are compared.)"""
    p, stars = synth.workload(133, LONG_R, 2)
    p.streams[1].parameters[4] = 1e-5
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh, timing=True) as ctx:
        got = ctx.eval(p.parameter_vector(), precision=sep.PREC_RESAMPLED)[0]
        tm = ctx.timing()
    ref = O.integrals(ParamsHolder(p), StarsHolder(stars))
    e = dict(I_bg=rel(got.background_integral, ref.es.background_integral),
             I_st=rel(got.stream_integrals[0], ref.stream_integrals()[0]))
    print({k: f"{v:.2e}" for k, v in e.items()}, "stream 2 integral", got.stream_integrals[1],
          "recurrence", tm.integral_points_recurrence, "of", tm.integral_points)
    assert got.stream_integrals[1] == 0.0
    assert tm.integral_points_recurrence == tm.integral_points > 0
    assert e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL


@pytest.mark.parametrize("sigma", [1.0, 0.3, 0.1])
def test_narrowest_streams_far_out(sigma):
    """Narrow streams for every stream, down to sigma = 0.1 kpc (the
    smallest stream_min of the reference's parameter files,
    tests/astronomy_parameters-11-small.txt; the synthetic workloads draw
    1-5 kpc)."""
    p, stars = synth.workload(97, LONG_R, 3)
    for s in p.streams:
        s.parameters[4] = sigma
    e, grew = errors_vs_oracle(p, stars)
    assert grew
    assert e["lnP"] <= LNP_TOL and e["I_st"] <= INT_REL_TOL and e["I_bg"] <= INT_REL_TOL


@pytest.mark.parametrize("q", [0.3, 1.0])
@pytest.mark.parametrize("r0", [1.0, 30.0])
@pytest.mark.parametrize("sigma", [0.1, 20.0])
def test_reference_bounds_against_oracle(q, r0, sigma):
    """The corners of the reference's parameter bounds (background_min /
    background_max: q 0.3-1.0, r0 1-30 kpc; stream_min / stream_max: sigma
    0.1-20 kpc, tests/astronomy_parameters-11-small.txt and -86-small.txt),
    every stream at sigma, on the full r and convolution resolution, against
    the oracle: every star and every integral within the bars."""
    p, stars = synth.workload(301, LONG_R, 3)
    p.background_parameters[1] = q
    p.background_parameters[2] = r0
    for s in p.streams:
        s.parameters[4] = sigma
    e, grew = errors_vs_oracle(p, stars)
    assert grew, "the resampled grid was not used"
    assert e["lnP"] <= LNP_TOL and e["lnL"] <= LNP_TOL
    assert e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL


@pytest.mark.parametrize("r", [130.0, 140.0])
def test_stream_far_outside_the_wedge(r):
    """A 1 kpc stream whose axis passes far outside the wedge (centre 130 or
    140 kpc out): every sight's terms lie 2^-480 to 2^-1000 below a sight
    through the stream, and the stream integral is 2^-481 or 2^-918, normal
    numbers in the oracle's FP64.  The
    integral still normalises the stream's per-star density, so it must
    match to the bars in every precision: the exact paths take each sight's
    minimum over the region's distances out of its exponents.  The resampled
    grids cannot resolve such steep tails (their fold would extrapolate the
    integrand past the region's ends), so a resampled request keeps the
    exact tables here: its integrals are the mixed path's bitwise."""
    p, stars = synth.workload(302, LONG_R, 1)
    st = p.streams[0].parameters
    st[1], st[2], st[3], st[4] = r, 1.047, 1.5708, 1.0
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    assert 0.0 < ref.stream_integrals[0] < 1e-140
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    got = {}
    with sep.Context(ph, sh, timing=True) as ctx:
        for prec, int_tol, lnp_tol in ((sep.PREC_FP64, 1e-12, 1e-10), (sep.PREC_MIXED, INT_REL_TOL, LNP_TOL),
                                       (sep.PREC_RESAMPLED, INT_REL_TOL, LNP_TOL)):
            g = got[prec] = ctx.eval(p.parameter_vector(), precision=prec, per_star=True)[0]
            tm = ctx.timing()
            e = dict(I_bg=rel(g.background_integral, ref.background_integral),
                     I_st=rel(g.stream_integrals, ref.stream_integrals),
                     lnP=math.log(10) * np.max(np.abs(g.log10_p - ref.log10_p)))
            print(r, prec, {k: f"{v:.2e}" for k, v in e.items()}, "I_st", g.stream_integrals[0],
                  "points", tm.integral_points)
            assert e["I_st"] <= int_tol and e["I_bg"] <= int_tol and e["lnP"] <= lnp_tol
    mixed, resampled = got[sep.PREC_MIXED], got[sep.PREC_RESAMPLED]
    assert resampled.background_integral == mixed.background_integral
    assert np.array_equal(resampled.stream_integrals, mixed.stream_integrals)


def test_stream_crossing_the_wedge_keeps_the_grid():
    """The same stream moved inside the wedge (centre 25 kpc out): the
    resampled grids are used, and the result is within the bars."""
    p, stars = synth.workload(302, LONG_R, 1)
    st = p.streams[0].parameters
    st[1], st[2], st[3], st[4] = 25.0, 1.047, 1.5708, 1.0
    e, grew = errors_vs_oracle(p, stars)
    assert grew, "the resampled grid was not used"
    assert e["lnP"] <= LNP_TOL and e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL


@pytest.mark.parametrize("field,value", [("sigma", 0.05), ("r0", 0.5), ("r0", 40.0), ("q", 0.2),
                                         ("q", 2.0)])
def test_outside_the_measured_domain_keeps_the_exact_tables(field, value):
    """Models outside the range the folds were measured over (sigma below
    0.1 kpc, r0 outside 1-30 kpc, q outside 0.3-1.6) run the exact tables
    in every region and the exact convolution in the star pass: the
    resampled request then gives the mixed path's result bitwise (for a
    narrow stream, whose width does not enter the star pass's background
    fold, the integrals bitwise and every star within the bar)."""
    p, stars = synth.workload(98, synth.WedgeShape(70, 24, 16, 30, 500), 2)
    if field == "sigma":
        p.streams[0].parameters[4] = value
    elif field == "r0":
        p.background_parameters[2] = value
    else:
        p.background_parameters[1] = value
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh, timing=True) as ctx:
        rs = ctx.eval(p.parameter_vector(), precision=sep.PREC_RESAMPLED, per_star=True)[0]
        tm = ctx.timing()
        mx = ctx.eval(p.parameter_vector(), precision=sep.PREC_MIXED, per_star=True)[0]
    assert tm.integral_points == sum(g.cells for g in p.regions) * p.convolve
    assert tm.integral_points_recurrence == 0
    assert rs.background_integral == mx.background_integral
    assert np.array_equal(rs.stream_integrals, mx.stream_integrals)
    if field == "sigma":
        assert math.log(10) * np.max(np.abs(rs.log10_p - mx.log10_p)) <= LNP_TOL
    else:
        assert rs.likelihood == mx.likelihood
        assert np.array_equal(rs.log10_p, mx.log10_p)


def test_sight_through_the_galactic_centre():
    """Stripe 52 near mu = 246 deg passes ~0.16 kpc from the Galactic centre:
    the background's sharpest peak, where the magnitude grid is the one
    that fits."""
    p, stars = synth.workload(98, LONG_R, 2)
    p.wedge = 52
    for reg in p.regions:
        reg.mu_min, reg.mu_max = 236.0, 256.0
    e, grew = errors_vs_oracle(p, stars)
    assert grew
    assert e["lnP"] <= LNP_TOL and e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL


def test_distance_grid_recurrence():
    """On wedges where a grid uniform in distance is chosen, the stream
    Gaussians come from the FP64 recurrence: the timing counters say so, and
    the stream integrals then match the oracle to the quadrature's own error
    (the fold of each stream's grid, below what an FP32 exponential per point
    gives)."""
    for seed, streams, sigma in ((121, 4, None), (122, 2, 1.0)):
        p, stars = synth.workload(seed, LONG_R, streams)
        if sigma:
            for st in p.streams:
                st.parameters[4] = sigma
        ph, sh = ParamsHolder(p), StarsHolder(stars)
        with sep.Context(ph, sh, timing=True) as ctx:
            got = ctx.eval(p.parameter_vector(), precision=sep.PREC_RESAMPLED, per_star=True)[0]
            tm = ctx.timing()
        ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
        e_st = rel(got.stream_integrals, ref.stream_integrals)
        lnp = math.log(10) * np.max(np.abs(got.log10_p - ref.log10_p))
        print(f"seed {seed}: {tm.integral_points_recurrence} of {tm.integral_points} by recurrence, "
              f"I_st {e_st:.2e} lnP {lnp:.2e}")
        assert tm.integral_points_recurrence > 0.5 * tm.integral_points
        assert e_st <= 3e-9   # the stream grids' fold (1.4e-9 at worst); FP32 per point gives ~1e-8
        assert lnp <= LNP_TOL


def background_stride_probe(ctx, vec):
    """(integral evaluations, background evaluations) of one resampled
    evaluation of vec."""
    ctx.eval(vec, precision=sep.PREC_RESAMPLED)
    tm = ctx.timing()
    return tm.integral_points, tm.integral_points_background


def test_background_grid():
    """Where the background is wider than the streams it runs on a grid of
    its own, uniform in magnitude: a fraction of the stream grid's nodes,
    and the background integral within the fold's own worst case (1.1e-10
    over the rule's whole parameter range; 2e-10 here, far inside the bar)
    of the oracle."""
    p, stars = synth.workload(91 + 3, LONG_R, 3)
    for st in p.streams:      # narrow streams: a fine distance grid
        st.parameters[4] = 1.0
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh, timing=True) as ctx:
        pts, bgp = background_stride_probe(ctx, p.parameter_vector())
        got = ctx.eval(p.parameter_vector(), precision=sep.PREC_RESAMPLED, per_star=True)[0]
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    e_bg = rel(got.background_integral, ref.background_integral)
    lnp = math.log(10) * np.max(np.abs(got.log10_p - ref.log10_p))
    print(f"{bgp} background evaluations for {pts} nodes; I_bg {e_bg:.2e}, lnP {lnp:.2e}")
    assert bgp <= 0.25 * pts
    assert e_bg <= 2e-10
    assert lnp <= LNP_TOL


def test_stream_grids():
    """With the background on its own grid each switched-on stream runs on a
    distance grid spaced for it alone: a 4 kpc stream beside a 1 kpc one
    costs fewer terms than the shared grid (nodes x streams), a switched-off
    stream none, and the results still match the oracle."""
    p, stars = synth.workload(95, LONG_R, 3)
    for st, sigma in zip(p.streams, (1.0, 4.0, 4.0)):
        st.parameters[4] = sigma
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    vec = p.parameter_vector()
    with sep.Context(ph, sh, timing=True) as ctx:
        got = ctx.eval(vec, precision=sep.PREC_RESAMPLED, per_star=True)[0]
        tm = ctx.timing()
        off = vec.copy()
        off[2 + 2 * 6 + 5] = 1e-5    # third stream's sigma: switched off
        ctx.eval(off, precision=sep.PREC_RESAMPLED)
        tm_off = ctx.timing()
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    e_st = rel(got.stream_integrals, ref.stream_integrals)
    lnp = math.log(10) * np.max(np.abs(got.log10_p - ref.log10_p))
    shared = tm.integral_points_recurrence * 3
    print(f"stream terms {tm.integral_stream_terms} (shared grid {shared}), with one off "
          f"{tm_off.integral_stream_terms}; I_st {e_st:.2e} lnP {lnp:.2e}")
    assert tm.integral_points_recurrence == tm.integral_points
    assert tm.integral_stream_terms < 0.7 * shared
    assert tm_off.integral_stream_terms < tm.integral_stream_terms
    assert e_st <= 1e-9 and lnp <= LNP_TOL


def test_background_stride(monkeypatch):
    """With the background's own grid switched off (MW_HIP_BG_GRID=0) it is
    sampled on every 2nd or 4th node of the distance grid where it is wider
    than the streams: fewer background evaluations than nodes, and the
    background integral still within the bar."""
    monkeypatch.setenv("MW_HIP_BG_GRID", "0")
    p, stars = synth.workload(91 + 3, LONG_R, 3)
    for st in p.streams:      # narrow streams: a fine grid, the background free to stride
        st.parameters[4] = 1.0
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh, timing=True) as ctx:
        pts, bgp = background_stride_probe(ctx, p.parameter_vector())
        got = ctx.eval(p.parameter_vector(), precision=sep.PREC_RESAMPLED)[0]
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    e_bg = rel(got.background_integral, ref.background_integral)
    print(f"{bgp} background evaluations for {pts} nodes; I_bg {e_bg:.2e}")
    assert 0.2 * pts <= bgp <= 0.55 * pts
    assert e_bg <= INT_REL_TOL


@pytest.mark.parametrize("bg_grid", ["1", "0"])
def test_graph_batch_replans_when_the_layout_changes(bg_grid, monkeypatch):
    """A graph-replayed batch given vectors whose tables differ only in how
    the background is sampled (same stream grid, different q: a different
    background grid, or with that grid off a different stride) must
    re-plan: each run equals the single evaluations of its own vectors."""
    monkeypatch.setenv("MW_HIP_BG_GRID", bg_grid)
    p, stars = synth.workload(131, LONG_R, 2)
    for st in p.streams:      # narrow streams: a fine grid, the background free to stride
        st.parameters[4] = 1.0
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    base = p.parameter_vector()
    with sep.Context(ph, sh, timing=True) as ctx:
        seen = {}
        for q in (0.45, 0.6, 0.8, 1.2, 1.6, 2.2, 3.0, 4.0):
            v = base.copy()
            v[0] = q
            seen[q] = (v, background_stride_probe(ctx, v))
        # same level (node counts within the stride tables' few extra nodes;
        # levels differ by sqrt 2), different background stride
        def stride(q):
            return round(seen[q][1][0] / seen[q][1][1])
        pairs = [(a, b) for a in seen for b in seen
                 if abs(seen[a][1][0] - seen[b][1][0]) <= 0.1 * seen[a][1][0] and
                 (stride(a) != stride(b) if bg_grid == "0" else seen[a][1][1] != seen[b][1][1])]
        assert pairs, f"no two q values share a level with different background sampling: {seen}"
        qa, qb = pairs[0]
        va, vb = seen[qa][0], seen[qb][0]
        print(f"q {qa} vs {qb}: {seen[qa][1]} / {seen[qb][1]} (nodes, background evaluations)")
        want_a = ctx.eval(va, precision=sep.PREC_RESAMPLED)[0]
        want_b = ctx.eval(vb, precision=sep.PREC_RESAMPLED)[0]
        with sep.Batch([(ctx, 1)], precision=sep.PREC_RESAMPLED, graph=True) as b:
            for vec, want in ((va, want_a), (vb, want_b), (va, want_a)):
                got = b.run([np.asarray([vec])])[0]
                assert same(got, want)


def aux_workload(seed, streams):
    """A wedge large enough to be resampled on a distance grid, with narrow
    streams (a fine grid) and the aux background profile (a, b, c = 100,
    50, 25, as the reference's aux fixtures)."""
    p, stars = synth.workload(seed, LONG_R, streams)
    for st in p.streams:
        st.parameters[4] = 1.0
    p.aux_bg_profile = 1
    p.background_parameters = list(p.background_parameters) + [100.0, 50.0, 25.0]
    p.background_step = list(p.background_step) + [1.0, 1.0, 1.0]
    p.background_min = list(p.background_min) + [-1000.0, -1000.0, -1000.0]
    p.background_max = list(p.background_max) + [1000.0, 1000.0, 1000.0]
    p.background_optimize = list(p.background_optimize) + [1, 1, 1]
    return p, stars


def test_aux_profile_resampled_single_and_batched():
    """The aux profile runs the FP64 kernels in every precision: resampled,
    its tables must stay in the (r, weight) layout (no recurrence, no
    background stride).  Single and batched evaluations agree bitwise and
    match the oracle."""
    p, stars = aux_workload(141, 2)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    vec = p.parameter_vector()
    with sep.Context(ph, sh, timing=True) as ctx:
        single = ctx.eval(vec, precision=sep.PREC_RESAMPLED, per_star=True)[0]
        tm = ctx.timing()
        with sep.Batch([(ctx, 1)], precision=sep.PREC_RESAMPLED, graph=False) as b:
            batched = b.run([np.asarray([vec])])[0]
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    e = dict(I_bg=rel(single.background_integral, ref.background_integral),
             I_st=rel(single.stream_integrals, ref.stream_integrals),
             lnP=math.log(10) * np.max(np.abs(single.log10_p - ref.log10_p)))
    print({k: f"{v:.2e}" for k, v in e.items()}, "points", tm.integral_points,
          "recurrence", tm.integral_points_recurrence, "background", tm.integral_points_background)
    assert tm.integral_points_recurrence == 0
    assert tm.integral_points_background == tm.integral_points
    assert same(batched, single)
    assert e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL and e["lnP"] <= LNP_TOL


def test_aux_profile_fixture():
    p, stars = fixture("86", n_stars=1500, aux=True)
    e, _ = errors_vs_oracle(p, stars)
    assert e["lnP"] <= LNP_TOL and e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL


def sweep_vectors(p, count, seed):
    """Seeded vectors over the reference's parameter range
    (background_min / background_max and stream_min / stream_max of its
    parameter files: q 0.3-1.0, r0 1-30 kpc, stream centres 0.7-57.5 kpc,
    sigma 0.1-20 kpc), q beyond it at 1.6, stream centres out to 60 kpc, and
    streams switched off (|sigma| < 1e-4)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    base = synth.parameter_vectors(seed % 1000, p, count)
    S = p.number_streams
    for j in range(count):
        v = base[j]
        v[0] = rng.choice([0.3, 0.45, 0.7, 0.9, 1.0, 1.6])
        v[1] = rng.choice([1.0, 2.0, 4.0, 8.0, 12.0, 20.0, 30.0])
        for s in range(S):
            o = 2 + 6 * s
            v[o + 2] = rng.choice([0.7, 2.0, 8.0, 25.0, 45.0, 57.5, 60.0])      # r
            v[o + 5] = rng.choice([0.1, 0.3, 0.6, 1.0, 2.5, 5.0, 12.0, 20.0])   # sigma
            if j % 11 == 5 and s == S - 1:
                v[o + 5] = 5e-5                                       # switched off
    return base


def test_sweep_against_gpu_fp64():
    """48 vectors on two wedges, each evaluated by the resampled path and by
    the GPU's FP64 path: every star within the bar and every integral within
    2e-8.  A switched-off stream's integral is 0 in both; its stars are
    compared only where the FP64 path's probability is finite."""
    worst = dict(I_bg=0.0, I_st=0.0, lnP=0.0)
    for k, shape in enumerate((LONG_R, synth.WedgeShape(350, 60, 24, 60, 1500))):
        p, stars = synth.workload(150 + k, shape, 3)
        vecs = sweep_vectors(p, 24, 7000 + k)
        ph, sh = ParamsHolder(p), StarsHolder(stars)
        with sep.Context(ph, sh) as ctx:
            got = ctx.eval(vecs, precision=sep.PREC_RESAMPLED, per_star=True)
            ref = ctx.eval(vecs, precision=sep.PREC_FP64, per_star=True)
        for g, r in zip(got, ref):
            worst["I_bg"] = max(worst["I_bg"], rel(g.background_integral, r.background_integral))
            live = r.stream_integrals != 0.0
            assert np.all(g.stream_integrals[~live] == 0.0)
            if live.any():
                worst["I_st"] = max(worst["I_st"], rel(g.stream_integrals[live], r.stream_integrals[live]))
            ok = np.isfinite(r.log10_p)
            assert np.array_equal(ok, np.isfinite(g.log10_p))
            if ok.any():
                worst["lnP"] = max(worst["lnP"], math.log(10) * float(np.max(np.abs(g.log10_p[ok] - r.log10_p[ok]))))
    print("worst over 48 vectors:", {k: f"{v:.2e}" for k, v in worst.items()})
    assert worst["lnP"] <= LNP_TOL
    assert worst["I_bg"] <= INT_REL_TOL and worst["I_st"] <= INT_REL_TOL


def test_sweep_over_stripes():
    """The same comparison on seven stripes, from low to high latitude and
    through the Galactic centre's direction (the magnitude grid's case),
    6 vectors each with the extremes above."""
    worst = dict(I_bg=0.0, I_st=0.0, lnP=0.0)
    for i, wedge in enumerate((10, 20, 37, 52, 61, 79, 86)):
        p, stars = synth.workload(160 + i, synth.WedgeShape(350, 60, 24, 60, 1500), 2)
        p.wedge = wedge
        vecs = sweep_vectors(p, 6, 8000 + i)
        ph, sh = ParamsHolder(p), StarsHolder(stars)
        with sep.Context(ph, sh) as ctx:
            got = ctx.eval(vecs, precision=sep.PREC_RESAMPLED, per_star=True)
            ref = ctx.eval(vecs, precision=sep.PREC_FP64, per_star=True)
        for g, r in zip(got, ref):
            e_bg = rel(g.background_integral, r.background_integral)
            live = r.stream_integrals != 0.0
            e_st = rel(g.stream_integrals[live], r.stream_integrals[live]) if live.any() else 0.0
            ok = np.isfinite(r.log10_p)
            assert np.array_equal(ok, np.isfinite(g.log10_p))
            lnp = math.log(10) * float(np.max(np.abs(g.log10_p[ok] - r.log10_p[ok]))) if ok.any() else 0.0
            worst = dict(I_bg=max(worst["I_bg"], e_bg), I_st=max(worst["I_st"], e_st), lnP=max(worst["lnP"], lnp))
        print(f"stripe {wedge}: worst so far", {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    assert worst["lnP"] <= LNP_TOL
    assert worst["I_bg"] <= INT_REL_TOL and worst["I_st"] <= INT_REL_TOL


def test_full_size_against_gpu_fp64():
    p, stars = synth.workload(0, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    print("full size:", len(stars), "stars", flush=True)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    vec = p.parameter_vector()
    with sep.Context(ph, sh, timing=True) as ctx:
        ref = ctx.eval(vec, precision=sep.PREC_FP64, per_star=True)[0]
        t64 = ctx.timing().integral_ms
        got = ctx.eval(vec, precision=sep.PREC_RESAMPLED, per_star=True)[0]
        trs = ctx.timing().integral_ms
    d = dict(I_bg=rel(got.background_integral, ref.background_integral),
             I_st=rel(got.stream_integrals, ref.stream_integrals),
             lnP=math.log(10) * np.max(np.abs(got.log10_p - ref.log10_p)))
    print({k: f"{v:.2e}" for k, v in d.items()}, f"integral ms: fp64 {t64:.1f} resampled {trs:.2f}")
    assert d["lnP"] <= LNP_TOL and d["I_bg"] <= INT_REL_TOL and d["I_st"] <= INT_REL_TOL
    assert trs < t64 / 10


SHAPES = [  # (shape, streams, vectors): two large enough to resample, one not
    (synth.WedgeShape(300, 20, 12, 60, 500), 2, 3),
    (synth.WedgeShape(500, 16, 16, 120, 400, cut=(500, 4, 16, 330.0, 340.0)), 4, 2),
    (synth.WedgeShape(20, 30, 16, 30, 300), 1, 2),
]


@pytest.fixture(scope="module")
def wedges():
    out = []
    for k, (shape, S, nv) in enumerate(SHAPES):
        p, stars = synth.workload(110 + k, shape, S)
        ph, sh = ParamsHolder(p), StarsHolder(stars)
        out.append((sep.Context(ph, sh), synth.parameter_vectors(110 + k, p, nv), ph, sh))
    yield out
    for ctx, *_ in out:
        ctx.close()


def same(a, b):
    return (a.likelihood == b.likelihood and a.background_integral == b.background_integral
            and np.array_equal(a.stream_integrals, b.stream_integrals) and a.num_zero == b.num_zero)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single(wedges, graph):
    singles = [ctx.eval(v, precision=sep.PREC_RESAMPLED)[0] for ctx, vecs, *_ in wedges for v in vecs]
    with sep.Batch([(c, len(v)) for c, v, *_ in wedges], precision=sep.PREC_RESAMPLED, graph=graph) as b:
        for _ in range(2 if graph else 1):
            got = b.run([v for _, v, *_ in wedges])
            assert all(same(g, r) for g, r in zip(got, singles))


@pytest.mark.parametrize("precision", [sep.PREC_MIXED, sep.PREC_RESAMPLED])
def test_batch_with_a_stream_outside_the_wedge(precision):
    """A graph-replayed batch whose middle vector puts its stream 140 kpc
    out, outside the wedge (the resampled plan keeps the exact tables for
    that vector alone): every result equals its single evaluation bitwise,
    on two replays, and the far vector's stream integral (8e-247) matches
    the oracle."""
    p, stars = synth.workload(302, LONG_R, 1)
    vecs = synth.parameter_vectors(302, p, 3)
    vecs[1, 4:8] = (140.0, 1.047, 1.5708, 1.0)      # r, theta, phi, sigma of the stream
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        singles = [ctx.eval(v, precision=precision)[0] for v in vecs]
        with sep.Batch([(ctx, 3)], precision=precision, graph=True) as b:
            for _ in range(2):
                got = b.run([vecs])
                assert all(same(g, r) for g, r in zip(got, singles))
    p.set_parameter_vector(vecs[1])
    ref = O.integrals(ParamsHolder(p), StarsHolder(stars))
    print("far stream integral", singles[1].stream_integrals[0], "oracle", ref.stream_integrals()[0])
    assert 0.0 < ref.stream_integrals()[0] < 1e-200
    assert rel(singles[1].stream_integrals, ref.stream_integrals()) <= INT_REL_TOL


@pytest.mark.parametrize("graph", [False, True])
def test_batch_mixing_split_and_strided_layouts(graph, monkeypatch):
    """One batch over a wedge in the split layout and a wedge whose
    background stays on the distance grid (built with MW_HIP_BG_GRID=0):
    the launch then runs the general persistent kernel instead of the
    split-only one, and every result equals its single evaluation bitwise;
    the split wedge alone runs the split-only kernel, with the same results."""
    p1, s1 = synth.workload(171, LONG_R, 3)
    p2, s2 = synth.workload(172, LONG_R, 2)
    for st in p2.streams:
        st.parameters[4] = 1.0
    c1 = sep.Context(ParamsHolder(p1), StarsHolder(s1), timing=True)
    monkeypatch.setenv("MW_HIP_BG_GRID", "0")
    c2 = sep.Context(ParamsHolder(p2), StarsHolder(s2), timing=True)
    monkeypatch.delenv("MW_HIP_BG_GRID")
    try:
        v1 = synth.parameter_vectors(171, p1, 3)
        v2 = synth.parameter_vectors(172, p2, 2)
        v2[:, 2 + 5::6] = 1.0                      # every stream 1 kpc wide: a fine grid
        single1 = [c1.eval(v, precision=sep.PREC_RESAMPLED)[0] for v in v1]
        t1 = c1.timing()
        single2 = [c2.eval(v, precision=sep.PREC_RESAMPLED)[0] for v in v2]
        t2 = c2.timing()
        # the first wedge's sights in the split layout (background on its own
        # grid), the second's with the background on the stream grid
        assert t1.integral_points_background < 0.5 * t1.integral_points
        assert t2.integral_points_background > 0.2 * t2.integral_points
        with sep.Batch([(c1, 3), (c2, 2)], precision=sep.PREC_RESAMPLED, graph=graph) as b:
            for _ in range(2 if graph else 1):
                got = b.run([v1, v2])
                assert all(same(g, r) for g, r in zip(got, single1 + single2))
        with sep.Batch([(c1, 3)], precision=sep.PREC_RESAMPLED, graph=graph) as b:
            got = b.run([v1])
            assert all(same(g, r) for g, r in zip(got, single1))
    finally:
        c1.close()
        c2.close()


def test_served_equals_single(wedges):
    with sep.Server(max_inflight=2, max_batch=4) as srv:
        ids = [srv.add_wedge(ph, sh) for _, _, ph, sh in wedges]
        rids, want = [], []
        for w, (ctx, vecs, *_) in enumerate(wedges):
            for v in vecs:
                rids.append(srv.submit(ids[w], v, sep.PREC_RESAMPLED))
                want.append(ctx.eval(v, precision=sep.PREC_RESAMPLED)[0])
        got = [srv.wait(r)[0] for r in rids]
    assert all(same(g, r) for g, r in zip(got, want))


def test_drop_in_does_not_take_resampled():
    with pytest.raises(sep.MwError):
        sep.set_precision(sep.PREC_RESAMPLED)


def test_narrow_stream_between_convolution_points():
    """A 0.5 kpc stream seen through 30 convolution points.  The mixed star
    pass rounds each stream exponent, shifted by its smallest value over the
    star's window, to FP32; where the Gaussian's peak lies inside the window
    between points more than 2 apart in t (p r_star above 2 / the widest
    ratio gap, 17.8 at 30 points: with k = 1.7 per kpc here, any star beyond
    ~11 kpc on a sight across the stream), every exponent there may be large,
    and the kernel splits them exactly instead.  Both paths run in this
    workload; every star stays within the bar."""
    shape = synth.WedgeShape(700, 24, 16, 30, 2000)
    p, stars = synth.workload(141, shape, 2)
    p.streams[0].parameters[4] = 0.5
    e, grew = errors_vs_oracle(p, stars)
    assert e["lnP"] <= LNP_TOL and e["lnL"] <= LNP_TOL
    assert e["I_bg"] <= INT_REL_TOL and e["I_st"] <= INT_REL_TOL
