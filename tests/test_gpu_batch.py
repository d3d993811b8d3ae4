"""Batched evaluation: many parameter vectors over wedges of uneven size and
different stream counts in one persistent launch (optionally a replayed HIP
graph).  Every vector's result must be bitwise the same as its single
evaluation, for every kernel variant, on every graph replay."""
import numpy as np
import pytest

from milkyway_client_amd import separation as sep
from milkyway_client_amd import synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder

pytestmark = pytest.mark.gpu

SHAPES = [  # (shape, streams, vectors)
    (synth.WedgeShape(20, 30, 16, 30, 700), 1, 3),
    (synth.WedgeShape(40, 50, 24, 60, 1500), 4, 2),
    (synth.WedgeShape(30, 64, 20, 120, 1000), 2, 4),
    (synth.WedgeShape(25, 40, 12, 20, 300, cut=(25, 10, 12, 330.0, 345.0)), 3, 2),
]


@pytest.fixture(scope="module")
def wedges():
    sep.load_library()
    sep.require_device()
    out = []
    for k, (shape, S, nv) in enumerate(SHAPES):
        p, stars = synth.workload(70 + k, shape, S)
        ph, sh = ParamsHolder(p), StarsHolder(stars)
        ctx = sep.Context(ph, sh)
        vecs = synth.parameter_vectors(70 + k, p, nv)
        out.append((ctx, vecs, ph, sh))
    yield out
    for ctx, *_ in out:
        ctx.close()


def singles(wedges, precision):
    ref = []
    for ctx, vecs, *_ in wedges:
        for v in vecs:
            ref.append(ctx.eval(v, precision=precision)[0])
    return ref


def same(a, b):
    return (a.likelihood == b.likelihood and a.background_integral == b.background_integral
            and np.array_equal(a.stream_integrals, b.stream_integrals)
            and a.num_zero == b.num_zero and a.prob_sum == b.prob_sum)


@pytest.mark.parametrize("precision", [sep.PREC_FP64, sep.PREC_MIXED])
def test_batch_equals_single_evaluations(wedges, precision):
    ref = singles(wedges, precision)
    with sep.Batch([(c, len(v)) for c, v, *_ in wedges], precision=precision, graph=False) as b:
        got = b.run([v for _, v, *_ in wedges])
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert same(g, r)


def test_every_variant_gives_the_same_bits(wedges):
    ref = singles(wedges, sep.PREC_MIXED)
    with sep.Batch([(c, len(v)) for c, v, *_ in wedges], precision=sep.PREC_MIXED, graph=False) as b:
        for variant in range(sep.Batch.variant_count()):
            for bpc in (1, 3):
                b.set_variant(variant, bpc)
                for g, r in zip(b.run([v for _, v, *_ in wedges]), ref):
                    assert same(g, r), (variant, bpc)


def test_graph_replays_match_eager(wedges):
    """The captured graph is replayed several times with new parameter values,
    with eager evaluations and synchronisations in between; every replay
    must equal the eager result."""
    groups = [(c, len(v)) for c, v, *_ in wedges]
    with sep.Batch(groups, precision=sep.PREC_FP64, graph=True) as g, \
            sep.Batch(groups, precision=sep.PREC_FP64, graph=False) as e:
        for rep in range(4):
            vecs = [synth.parameter_vectors(80 + rep + 10 * k, ph.py, len(v))
                    for k, (_, v, ph, _) in enumerate(wedges)]
            want = e.run(vecs)
            wedges[0][0].eval(wedges[0][1][0])          # unrelated eager work
            sep.synchronize()
            got = g.run(vecs)
            for a, b in zip(got, want):
                assert same(a, b), rep


def test_batch_rejects_bad_input(wedges):
    ctx = wedges[0][0]
    with pytest.raises(sep.MwError):
        sep.Batch([(ctx, 0)])
    with sep.Batch([(ctx, 1)]) as b:
        with pytest.raises(sep.MwError):
            b.set_variant(99, 1)


def test_variant_scratch_query(wedges):
    """mw_batch_variant_scratch reports each variant's spill bytes once the
    batch has planned its launch (its first run), and refuses before."""
    groups = [(c, len(v)) for c, v, *_ in wedges]
    with sep.Batch(groups, precision=sep.PREC_MIXED, graph=False) as b:
        with pytest.raises(sep.MwError):
            b.variant_scratch(0)
        b.run([v for _, v, *_ in wedges])
        sizes = [b.variant_scratch(v) for v in range(sep.Batch.variant_count())]
        assert all(isinstance(s, int) and 0 <= s < 1 << 20 for s in sizes)
        with pytest.raises(sep.MwError):
            b.variant_scratch(sep.Batch.variant_count())
