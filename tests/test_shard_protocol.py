"""The sharded-evaluation protocol (milkyway_client_amd.shard) across real
processes over gloo, on the CPU: 2 and 3 ranks, each started as its own
process, split the reference's parameter file 86 (a main region and a cut)
and its first 2000 stars as mpi_synchronous_astronomy.c:45-83 does, gather
their stand-in integral_f results, compose, gather their stand-in
likelihood_f results and compose (astronomy_worker.c:68-86, :115-134).

Checks: the ranks' calculation ranges partition every region and their star
shares partition the stars; every rank's composed integrals are the whole
wedge's cell count and index sum, and those are what each rank's
likelihood_f received; every rank returns the same likelihood, bitwise, equal
to the mean of l + b + r over all stars minus 3.
"""
from pathlib import Path

import numpy as np
import pytest

from helpers import GOLDEN, run_ranks
from milkyway_client_amd import params_io

WORKER = Path(__file__).with_name("shard_protocol_worker.py")
PARAMS = GOLDEN / "astronomy_parameters-86-small.txt"
STARS = GOLDEN / "stars-86-first2000.txt"


@pytest.mark.parametrize("world", [2, 3])
def test_protocol_composes_over_gloo(world):
    outs = run_ranks([str(WORKER), str(PARAMS), str(STARS)], world)
    p = params_io.read_astronomy_parameters(PARAMS)
    totals = [g.r_steps * g.mu_steps * g.nu_steps for g in p.regions]
    stars = np.loadtxt(STARS, skiprows=1)
    assert [o["rank"] for o in outs] == list(range(world))
    # the ranges partition every region, the star shares partition the stars
    for g, total in enumerate(totals):
        edges = [o["ranges"][g] for o in outs]
        assert edges[0][0] == 0 and edges[-1][1] == total
        assert all(edges[r][1] == edges[r + 1][0] for r in range(world - 1))
    assert sum(o["stars"] for o in outs) == len(stars) == outs[0]["total_stars"]
    # composed integrals: the whole wedge's counts, on every rank
    want = [float(sum(totals)), float(sum(t * (t - 1) // 2 for t in totals))]
    want += [float(sum(g * t for g, t in enumerate(totals)))] * (p.number_streams - 1)
    for o in outs:
        assert o["integrals"] == want
        assert o["received_integrals"] == want
    # every rank the same likelihood, bitwise; the value from all the stars
    assert len({o["likelihood"] for o in outs}) == 1
    expect = float(np.sum(stars[:, 0] + stars[:, 1] + stars[:, 2])) / len(stars) - 3.0
    assert abs(outs[0]["likelihood"] - expect) <= 1e-12 * abs(expect)
