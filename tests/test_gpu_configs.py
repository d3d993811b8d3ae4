"""The benchmark's workloads at their stated sizes, each checked against an
independent result.

* Full size, FP64 (700 x 1600 x 640 cells plus a 700 x 400 x 160 cut,
  n = 120, 3 streams, 200k stars): the whole evaluation against the oracle's
  committed result for the same inputs (tests/golden/full_size_oracle.json,
  written by tests/golden/make_full_size_fixture.py): region integrals within
  1e-12 relative, the likelihood within 1e-10, every tenth star's log10 P
  within 1e-10, and the sum of all 200k log10 P.
  Reference: evaluation_optimized.c:424-753.
* Full size, mixed precision: 32 seeded parameter vectors against the FP64
  path on the same shape (itself pinned to the oracle by the test above):
  every star's |d ln P| <= 1e-7.
* The batch: 16 parameter vectors (4 streams) on each of the 4 uneven wedges
  70x160x64 (n = 30, 10k stars), 350x400x160 (n = 60, 50k),
  700x800x320 (n = 120, 100k) and 700x1600x640 (n = 120, 200k), r x mu x nu,
  in one persistent launch replayed from a HIP graph, for the exact and the
  resampled quadrature: every vector bitwise equal to its single evaluation,
  and every star within 1e-7 in ln P of the FP64 path.
* Serving: a Poisson arrival stream over 8 resident wedges drawn from the
  batch shapes (1-4 streams, exact mixed, resampled and FP64 requests): every
  result bitwise equal to the same request evaluated alone on a context.
"""
import json
import math

import numpy as np
import pytest

from helpers import GOLDEN, rel
from milkyway_client_amd import separation as sep
from milkyway_client_amd import serving, synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder

pytestmark = pytest.mark.gpu

INT_TOL = 1e-12      # FP64 integrals, relative
LNL_TOL = 1e-10      # FP64 likelihood and per-star log10 P
LNP_TOL = 1e-7       # mixed precision, per star, natural log
LN10 = math.log(10.0)


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


@pytest.fixture(scope="module")
def full_size():
    p, stars = synth.workload(1, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    ctx = sep.Context(ph, sh)
    yield p, stars, ctx
    ctx.close()


def same(a, b):
    return (a.likelihood == b.likelihood and a.prob_sum == b.prob_sum
            and a.background_integral == b.background_integral
            and np.array_equal(a.stream_integrals, b.stream_integrals) and a.num_zero == b.num_zero)


def test_full_size_fp64_against_the_oracle(full_size):
    p, stars, ctx = full_size
    ref = json.loads((GOLDEN / "full_size_oracle.json").read_text())
    ref_lp = np.load(GOLDEN / "full_size_log10p_every10.npy")
    vec = np.asarray(ref["parameter_vector"])
    assert np.array_equal(vec, p.parameter_vector())
    got = ctx.eval(vec, precision=sep.PREC_FP64, per_star=True, regions=True)[0]
    S = p.number_streams
    for g, row in enumerate(ref["regions"]):
        assert rel(got.regions[g][:1 + S], row) <= INT_TOL, g
    assert rel(got.background_integral, ref["background_integral"]) <= INT_TOL
    assert rel(got.stream_integrals, ref["stream_integrals"]) <= INT_TOL
    assert got.num_zero == ref["num_zero"]
    assert abs(got.likelihood - ref["likelihood"]) <= LNL_TOL
    assert abs(got.bg_only_likelihood - ref["bg_only_mean"]) <= LNL_TOL
    assert np.max(np.abs(got.stream_only_likelihood - np.asarray(ref["st_only_mean"]))) <= LNL_TOL
    lp = got.log10_p
    assert lp.size == ref["n_stars"]
    assert np.max(np.abs(lp[::ref["log10p_stride"]] - ref_lp)) <= LNL_TOL
    assert abs(float(np.sum(lp)) - ref["sum_log10p_all_stars"]) / lp.size <= LNL_TOL


def test_full_size_mixed_32_vectors(full_size):
    p, stars, ctx = full_size
    vecs = synth.parameter_vectors(2, p, 32)
    worst = {"lnP": 0.0, "I_bg": 0.0, "I_st": 0.0, "lnL": 0.0}
    for k in range(0, 32, 8):
        rows = vecs[k:k + 8]
        ref = ctx.eval(rows, precision=sep.PREC_FP64, per_star=True)
        got = ctx.eval(rows, precision=sep.PREC_MIXED, per_star=True)
        for g, r in zip(got, ref):
            worst["lnP"] = max(worst["lnP"], LN10 * float(np.max(np.abs(g.log10_p - r.log10_p))))
            worst["I_bg"] = max(worst["I_bg"], rel(g.background_integral, r.background_integral))
            worst["I_st"] = max(worst["I_st"], rel(g.stream_integrals, r.stream_integrals))
            worst["lnL"] = max(worst["lnL"], LN10 * abs(g.likelihood - r.likelihood))
    print("mixed vs FP64, full size, 32 vectors:", {k: f"{v:.2e}" for k, v in worst.items()})
    assert worst["lnP"] <= LNP_TOL
    assert worst["lnL"] <= LNP_TOL


@pytest.fixture(scope="module")
def batch_wedges():
    out = []
    for k, shape in enumerate(synth.BATCH_SHAPES):
        p, stars = synth.workload(100 + k, shape, 4)
        ctx = sep.Context(ParamsHolder(p), StarsHolder(stars))
        out.append((ctx, synth.parameter_vectors(100 + k, p, 16)))
    yield out
    for ctx, _ in out:
        ctx.close()


@pytest.mark.parametrize("precision", [sep.PREC_MIXED, sep.PREC_RESAMPLED])
def test_batch_workload(batch_wedges, precision):
    groups = [(c, len(v)) for c, v in batch_wedges]
    with sep.Batch(groups, precision=precision, graph=True) as b:
        b.run([v for _, v in batch_wedges])            # captures the graph
        got = b.run([v for _, v in batch_wedges])      # replays it
    assert len(got) == 64
    i = 0
    worst = 0.0
    for ctx, vecs in batch_wedges:
        single = ctx.eval(vecs, precision=precision, per_star=True)
        ref = ctx.eval(vecs, precision=sep.PREC_FP64, per_star=True)
        for s, r in zip(single, ref):
            assert same(got[i], s), i
            worst = max(worst, LN10 * float(np.max(np.abs(s.log10_p - r.log10_p))))
            i += 1
    print(f"batch precision {precision}: worst per-star |d ln P| {worst:.2e} over 64 vectors")
    assert worst <= LNP_TOL


def test_serving_workload():
    wedges = serving.make_wedges(8, seed=4000)
    holders = [(ParamsHolder(p), StarsHolder(s)) for p, s, _ in wedges]
    reqs = serving.make_requests(wedges, 24, seed=5000, fast_precision=sep.PREC_RESAMPLED)
    reqs += serving.make_requests(wedges, 24, seed=5100, fast_precision=sep.PREC_MIXED,
                                  fp64_fraction=0.0)
    order = np.random.Generator(np.random.PCG64(5200)).permutation(len(reqs))
    reqs = [reqs[i] for i in order]
    arrivals = serving.poisson_arrivals(len(reqs), rate=100.0, seed=5300)
    with sep.Server(max_inflight=3, max_batch=16) as srv:
        ids = [srv.add_wedge(ph, sh) for ph, sh in holders]
        out = serving.run(srv, ids, reqs, arrivals)
        st = srv.stats()
    assert out["n"] == len(reqs) and st.completed == len(reqs)
    ctxs = [sep.Context(ph, sh) for ph, sh in holders]
    try:
        for r, got in zip(reqs, out["results"]):
            alone = ctxs[r.wedge].eval(r.vector, precision=r.precision)[0]
            assert same(alone, got)
    finally:
        for c in ctxs:
            c.close()
