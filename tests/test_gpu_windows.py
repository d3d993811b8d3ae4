"""Regions of few 256-column tiles are split along r.

A region with fewer than about 2,048 tiles gets r_chunks windows of its
r-steps per tile (at least 4 r-steps each), so that a launch of the region
alone fills the GPU; the count depends on the region alone.  These tests
check the launch's block count against that rule and the result against the
oracle at the FP64 bars (integrals 1e-12 relative, every star's log10 P and
the likelihood 1e-10), for wedges of 1, 40 and 250 tiles, and that mixed
precision stays within its bar on the same windows.
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

TARGET_BLOCKS, MIN_R_STEPS, TILE = 2048, 4, 256


def windows(tiles, r_steps):
    want = (TARGET_BLOCKS + tiles - 1) // tiles
    return max(1, min(want, r_steps // MIN_R_STEPS))


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


@pytest.mark.parametrize("shape", [synth.WedgeShape(13, 16, 16, 20, 500),       # 1 tile, 3 windows
                                   synth.WedgeShape(70, 160, 64, 30, 2000),     # 40 tiles, 17 windows
                                   synth.WedgeShape(40, 400, 160, 12, 2000)])   # 250 tiles, 9 windows
def test_windows_block_count_and_parity(shape):
    p, stars = synth.workload(21, shape, 2)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    tiles = [math.ceil(g.mu_steps * g.nu_steps / TILE) for g in p.regions]
    expect = sum(t * windows(t, g.r_steps) for t, g in zip(tiles, p.regions))
    ref = O.evaluate_full(ph, sh)
    with sep.Context(ph, sh, timing=True) as ctx:
        got = ctx.eval(p.parameter_vector(), precision=sep.PREC_FP64, per_star=True)[0]
        blocks = ctx.timing().integral_blocks
        mixed = ctx.eval(p.parameter_vector(), precision=sep.PREC_MIXED, per_star=True)[0]
    print(shape, "blocks", blocks, "expected", expect)
    assert blocks == expect
    assert rel(got.background_integral, ref.background_integral) <= 1e-12
    assert rel(got.stream_integrals, ref.stream_integrals) <= 1e-12
    assert np.max(np.abs(got.log10_p - ref.log10_p)) <= 1e-10
    assert abs(got.likelihood - ref.likelihood) <= 1e-10
    assert math.log(10) * np.max(np.abs(mixed.log10_p - ref.log10_p)) <= 1e-7
