"""One rank of the host-only test of the sharded-evaluation protocol
(milkyway_client_amd.shard), started once per rank by
tests/test_shard_protocol.py with RANK, WORLD_SIZE, MASTER_ADDR and
MASTER_PORT set.

The rank reads the parameter and star files, splits them as the reference's
MPI driver does (mpi_synchronous_astronomy.c:45-83), and runs the protocol
with stand-in plugins whose composed values have closed forms:
  integral_f   -> [cells in the rank's ranges, sum of their flat indices,
                   sum over regions of region index x cells, ...]
  likelihood_f -> [sum over the rank's stars of l + b + r, 0]
so the composed integrals are the whole wedge's counts and the likelihood is
the mean of l + b + r over all stars minus 3.  It prints one JSON line.
"""
import datetime
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import torch.distributed as dist  # noqa: E402

from milkyway_client_amd.shard import ShardedEvaluator  # noqa: E402


def cells_integral_f(ev, vector, results):
    results[:] = 0.0
    for g in range(ev.es.number_integrals):
        a = ev.es.integral[g].contents
        lo, hi = a.min_calculation, a.max_calculation
        results[0] += hi - lo
        if results.size > 1:
            results[1] += (lo + hi - 1) * (hi - lo) // 2
        for s in range(2, results.size):
            results[s] += g * (hi - lo)


def stars_likelihood_f(ev, integrals, results):
    ev.received_integrals = integrals.copy()
    total = 0.0
    for i in range(ev.sp.number_stars):
        row = ev.sp.stars[i]
        total += row[0] + row[1] + row[2]
    results[0] = total
    results[1] = 0.0


def main():
    params, stars = sys.argv[1], sys.argv[2]
    dist.init_process_group("gloo", timeout=datetime.timedelta(seconds=120))
    try:
        with ShardedEvaluator(params, stars, integral_f=cells_integral_f,
                              likelihood_f=stars_likelihood_f) as ev:
            lnl = ev.evaluate(np.zeros(ev.n_params))
            ranges = [(ev.es.integral[g].contents.min_calculation, ev.es.integral[g].contents.max_calculation)
                      for g in range(ev.es.number_integrals)]
            out = {"rank": ev.rank, "world": ev.world, "likelihood": lnl,
                   "integrals": ev.last_integrals.tolist(),
                   "received_integrals": ev.received_integrals.tolist(),
                   "ranges": ranges, "stars": ev.sp.number_stars, "total_stars": ev.total_stars}
        print(json.dumps(out), flush=True)
    finally:
        dist.destroy_process_group()


if __name__ == "__main__":
    main()
