"""Request server: continuous batching over resident wedges.

Each request's result must be bitwise the same alone and in the mix (the
integral and star sums of a request are reduced in a fixed order of its own
blocks), requests admitted under a tight HBM budget must still complete, and
the budget must refuse wedges that do not fit."""
import numpy as np
import pytest

from milkyway_client_amd import separation as sep
from milkyway_client_amd import serving, synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder

pytestmark = pytest.mark.gpu

SHAPES = [synth.WedgeShape(20, 30, 16, 30, 800), synth.WedgeShape(40, 60, 24, 60, 1500),
          synth.WedgeShape(30, 40, 20, 20, 300, cut=(30, 10, 20, 330.0, 345.0))]


@pytest.fixture(scope="module")
def wedges():
    sep.load_library()
    sep.require_device()
    return serving.make_wedges(5, seed=500, shapes=SHAPES)


def same(a, b):
    return (a.likelihood == b.likelihood and a.background_integral == b.background_integral
            and np.array_equal(a.stream_integrals, b.stream_integrals) and a.num_zero == b.num_zero)


def test_same_alone_and_in_the_mix(wedges):
    reqs = serving.make_requests(wedges, 24, seed=600, fp64_fraction=0.3)
    holders = [(ParamsHolder(p), StarsHolder(s)) for p, s, _ in wedges]
    with sep.Server(max_inflight=1, max_batch=1) as alone:
        ids = [alone.add_wedge(ph, sh) for ph, sh in holders]
        solo = [alone.wait(alone.submit(ids[r.wedge], r.vector, r.precision))[0] for r in reqs]
    with sep.Server(max_inflight=3, max_batch=8) as mix:
        ids = [mix.add_wedge(ph, sh) for ph, sh in holders]
        rids = [mix.submit(ids[r.wedge], r.vector, r.precision) for r in reqs]
        mixed = [mix.wait(rid)[0] for rid in rids]
        st = mix.stats()
    assert st.completed == len(reqs) and st.batches < len(reqs)
    for a, b in zip(solo, mixed):
        assert same(a, b)
    # and equal to a plain single evaluation on a context
    ctxs = [sep.Context(ph, sh) for ph, sh in holders]
    for r, a in zip(reqs[:6], solo):
        assert same(ctxs[r.wedge].eval(r.vector, precision=r.precision)[0], a)


def test_poisson_stream_completes(wedges):
    reqs = serving.make_requests(wedges, 30, seed=700)
    arrivals = serving.poisson_arrivals(len(reqs), rate=200.0, seed=701)
    with sep.Server(max_inflight=3, max_batch=16) as srv:
        ids = [srv.add_wedge(ParamsHolder(p), StarsHolder(s)) for p, s, _ in wedges]
        out = serving.run(srv, ids, reqs, arrivals)
    assert out["n"] == len(reqs)
    assert out["p99_ms"] >= out["p50_ms"] > 0
    assert all(np.isfinite(r.likelihood) for r in out["results"])


def test_hbm_budget_admission(wedges):
    p, s, _ = wedges[0]
    ph, sh = ParamsHolder(p), StarsHolder(s)
    with sep.Context(ph, sh) as ctx:
        resident = ctx.resident_bytes
    # room for the wedge and about one request's scratch (~1.6 KB) at a time
    budget = resident + 2_500
    with sep.Server(budget_bytes=budget, max_inflight=2, max_batch=16) as srv:
        wid = srv.add_wedge(ph, sh)
        with pytest.raises(sep.MwError):
            srv.add_wedge(ph, sh)                 # a second copy does not fit
        vecs = synth.parameter_vectors(800, p, 12)
        rids = [srv.submit(wid, v) for v in vecs]
        res = [srv.wait(r)[0] for r in rids]
        st = srv.stats()
    assert st.completed == 12 and st.largest_batch < 12
    single = sep.Context(ph, sh)
    for v, r in zip(vecs[:3], res):
        assert same(single.eval(v, precision=sep.PREC_MIXED)[0], r)


def test_request_collection(wedges):
    p, s, _ = wedges[0]
    with sep.Server(max_inflight=1, max_batch=4) as srv:
        wid = srv.add_wedge(ParamsHolder(p), StarsHolder(s))
        rid = srv.submit(wid, p.parameter_vector())
        srv.wait(rid)
        with pytest.raises(sep.MwError):
            srv.wait(rid)                          # already collected
        with pytest.raises(sep.MwError):
            srv.poll(123456)
        with pytest.raises(sep.MwError):
            srv.submit(99, p.parameter_vector())   # unknown wedge
