"""The autotuner's choice among timed kernel variants (autotune.pick): the
fastest variant whose kernels spill no registers, unless a spilling one is
more than the margin faster."""
from milkyway_client_amd import autotune


TABLE = [(0, 4, 130.0), (1, 4, 130.3), (3, 4, 128.9), (5, 4, 128.1), (5, 5, 129.5)]


def test_prefers_a_clean_variant_within_the_margin():
    spills = {0: 0, 1: 0, 3: 0, 5: 116}
    assert autotune.pick(TABLE, spills) == (3, 4, 128.9)


def test_takes_the_fastest_when_no_clean_variant_is_close():
    spills = {0: 0, 1: 0, 3: 36, 5: 116}
    assert autotune.pick(TABLE, spills) == (5, 4, 128.1)          # 130.0 is 1.5 % slower
    assert autotune.pick(TABLE, spills, margin=0.02) == (0, 4, 130.0)


def test_every_variant_spilling_gives_the_fastest():
    spills = {0: 8, 1: 8, 3: 36, 5: 116}
    assert autotune.pick(TABLE, spills) == (5, 4, 128.1)


def test_fastest_clean_variant_wins_outright():
    spills = {0: 0, 1: 0, 3: 0, 5: 0}
    assert autotune.pick(TABLE, spills) == (5, 4, 128.1)
