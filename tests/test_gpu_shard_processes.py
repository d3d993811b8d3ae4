"""Sharded evaluation in separate processes on the GPU: one process per rank
(milkyway_client_amd.shard, the reference's MPI driver protocol of
mpi_synchronous_astronomy.c:45-172 over torch.distributed gloo), all ranks
sharing the one GPU, each running the library's integral_f and likelihood_f
plugins on its share of the reference's parameter file and stars.  Every
rank's composed integrals must match the oracle's within 1e-12 relative and
its likelihood the oracle's evaluate() - 3.0 within 1e-10 in FP64 (mixed
precision: 1e-8 and 1e-7 / ln 10), and every rank must return the same
values, bitwise.
"""
import pytest

from helpers import GOLDEN, ctest_vectors, fixture, rel, run_ranks
from milkyway_client_amd import separation as sep
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()


# (integrals relative, likelihood absolute) per precision: FP64 at the
# parity bars; mixed at its per-star bar (the likelihood is a mean of
# log10 P, so |d lnL| <= 1e-7 bounds it by 1e-7 / ln 10) and 1e-8 on integrals
BARS = {"fp64": (1e-12, 1e-10), "mixed": (1e-8, 1e-7 / 2.302585092994046)}


# A rank's first `import torch` on a fresh machine can take a minute or two.
@pytest.mark.timeout(480)
@pytest.mark.parametrize("fid,world,precision", [("86", 2, "fp64"), ("11", 3, "fp64"), ("79", 2, "mixed")])
def test_ranks_in_processes_compose(fid, world, precision):
    vec = ctest_vectors()[fid]
    cmd = ["-m", "milkyway_client_amd.shard",
           "--params", str(GOLDEN / f"astronomy_parameters-{fid}-small.txt"),
           "--stars", str(GOLDEN / f"stars-{fid}-first2000.txt"),
           "--precision", precision,
           "--vector"] + [repr(float(x)) for x in vec]
    outs = run_ranks(cmd, world, timeout=420)
    int_tol, lnl_tol = BARS[precision]
    p, stars = fixture(fid, n_stars=2000)
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    print(fid, world, [o["likelihood"] for o in outs], ref.likelihood - 3.0)
    assert [o["rank"] for o in outs] == list(range(world))
    assert sum(o["stars"] for o in outs) == len(stars)
    for o in outs:
        assert (o["likelihood"], o["background_integral"], o["stream_integrals"]) == \
            (outs[0]["likelihood"], outs[0]["background_integral"], outs[0]["stream_integrals"])
    assert rel(outs[0]["background_integral"], ref.background_integral) <= int_tol
    assert rel(outs[0]["stream_integrals"], ref.stream_integrals) <= int_tol
    assert abs(outs[0]["likelihood"] - (ref.likelihood - 3.0)) <= lnl_tol
