"""Sharded evaluation: one process per GPU, the reference's distributed
evaluator protocol over torch.distributed.

The reference's MPI search (milkyway_separation/cpu/mpi_synchronous_astronomy.c:45-172,
with the plugin signatures of astronomy_worker.c:35-134) has every rank read
the full parameter and star files, keep its share of each region's flat
calculation range (split_astronomy_parameters, parameters.c:502-522) and its
contiguous share of the stars (split_star_points, star_points.c:106-130), and
then, per parameter vector:

1. integral_f: the integrals over the rank's ranges (1 + S doubles);
2. integral_compose: the ranks' results summed in rank order;
3. likelihood_f: the star pass over the rank's stars with the composed
   integrals (prob_sum and bad_jacobians);
4. likelihood_compose: the ranks' sums over the total star count, minus 3.

Here steps 1 and 3 are the library's GPU plugins (mw_hip_integral_f,
mw_hip_likelihood_f) on the rank's GPU, steps 2 and 4 its host compose
functions, and the exchange between them an all-gather of those few doubles
over the process group (gloo: host tensors, so no device memory crosses
processes).  Every rank composes the same gathered values in rank order, so
every rank returns the same likelihood, bitwise.

Run one process per GPU (torchrun or any launcher that sets RANK,
WORLD_SIZE, LOCAL_RANK, MASTER_ADDR and MASTER_PORT):

    python -m milkyway_client_amd.shard --params astronomy_parameters.txt \\
        --stars stars.txt --vector v0 v1 ... [--json out.json]
"""
from __future__ import annotations

import argparse
import ctypes as C
import datetime
import json
import os
import sys
from typing import Callable, Optional

import numpy as np

from . import separation as sep
from .refstructs import ASTRONOMY_PARAMETERS, EVALUATION_STATE, STAR_POINTS

_dp = C.POINTER(C.c_double)

# A rank's integral_f or likelihood_f (astronomy_worker.c:35-66, :88-113):
# called as fn(evaluator, inputs, results) with the evaluator's split ap, es
# and sp, the parameter vector (or the composed integrals) in inputs, and
# results (1 + S doubles, or prob_sum and bad_jacobians) written in place.
PluginFn = Callable[["ShardedEvaluator", np.ndarray, np.ndarray], None]


def gpu_integral_f(ev: "ShardedEvaluator", vector: np.ndarray, results: np.ndarray) -> None:
    ev._L.mw_hip_integral_f(vector.ctypes.data_as(_dp), results.ctypes.data_as(_dp), C.byref(ev.ap),
                            C.byref(ev.es), C.byref(ev.sp))


def gpu_likelihood_f(ev: "ShardedEvaluator", integrals: np.ndarray, results: np.ndarray) -> None:
    ev._L.mw_hip_likelihood_f(integrals.ctypes.data_as(_dp), results.ctypes.data_as(_dp), C.byref(ev.ap),
                              C.byref(ev.es), C.byref(ev.sp))


class ShardedEvaluator:
    """The calling rank's share of a wedge and its part of the protocol.

    The process group must be initialised (torch.distributed) before
    construction.  integral_f and likelihood_f default to the library's GPU
    plugins; a caller may pass others with the same arguments (the host-only
    test of the protocol does)."""

    def __init__(self, params_file: str, stars_file: str, integral_f: Optional[PluginFn] = None,
                 likelihood_f: Optional[PluginFn] = None, group=None):
        import torch.distributed as dist
        self._dist = dist
        self._group = group
        self.rank = dist.get_rank(group)
        self.world = dist.get_world_size(group)
        L = sep.load_library()
        self._L = L
        self.ap, self.sp, self.es = ASTRONOMY_PARAMETERS(), STAR_POINTS(), EVALUATION_STATE()
        self._owned = []
        if L.mw_read_astronomy_parameters(os.fsencode(params_file), C.byref(self.ap)) != 0:
            raise sep.MwError(sep.MW_E_INVALID, f"cannot read {params_file}")
        self._owned.append(lambda: L.mw_free_astronomy_parameters(C.byref(self.ap)))
        if L.mw_read_star_points(os.fsencode(stars_file), C.byref(self.sp)) != 0:
            raise sep.MwError(sep.MW_E_INVALID, f"cannot read {stars_file}")
        self._owned.append(lambda: L.mw_free_star_points(C.byref(self.sp)))
        self.total_stars = self.sp.number_stars
        L.mw_split_astronomy_parameters(C.byref(self.ap), self.rank, self.world)
        if L.mw_split_star_points(C.byref(self.sp), self.rank, self.world) != 0:
            raise sep.MwError(sep.MW_E_NOMEM, "split_star_points failed")
        if L.mw_initialize_state(C.byref(self.ap), C.byref(self.sp), C.byref(self.es)) != 0:
            raise sep.MwError(sep.MW_E_NOMEM, "initialize_state failed")
        self._owned.append(lambda: L.mw_free_state(C.byref(self.es)))
        self.n_params = L.mw_get_optimized_parameter_count(C.byref(self.ap))
        self.S = self.ap.number_streams
        self.integral_f = integral_f or gpu_integral_f
        self.likelihood_f = likelihood_f or gpu_likelihood_f
        self.last_integrals = None

    def _gather(self, values: np.ndarray) -> np.ndarray:
        """Every rank's values, stacked in rank order (host tensors)."""
        import torch
        t = torch.from_numpy(np.ascontiguousarray(values, dtype=np.float64).copy())
        parts = [torch.empty_like(t) for _ in range(self.world)]
        self._dist.all_gather(parts, t, group=self._group)
        return np.ascontiguousarray(torch.stack(parts).numpy())

    def integrals(self, vector) -> np.ndarray:
        """Steps 1-2: the composed background and stream integrals."""
        v = np.ascontiguousarray(vector, dtype=np.float64)
        if v.size != self.n_params:
            raise ValueError(f"parameter vectors need {self.n_params} values, got {v.size}")
        part = np.zeros(1 + self.S)
        self.integral_f(self, v, part)
        every = self._gather(part)
        composed = np.zeros(1 + self.S)
        self._L.mw_integral_compose(every.ctypes.data_as(_dp), self.world, composed.ctypes.data_as(_dp),
                                    C.byref(self.ap))
        return composed

    def evaluate(self, vector) -> float:
        """The client's likelihood of one parameter vector (mean log10 P - 3);
        the composed integrals it used are kept in last_integrals."""
        composed = self.integrals(vector)
        self.last_integrals = composed
        part = np.zeros(2)
        self.likelihood_f(self, composed, part)
        every = self._gather(part)
        return float(self._L.mw_likelihood_compose(every.ctypes.data_as(_dp), self.world, self.total_stars))

    def close(self):
        while self._owned:
            self._owned.pop()()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--params", required=True, help="astronomy parameter file (reference format)")
    ap.add_argument("--stars", required=True, help="star file (reference format)")
    ap.add_argument("--vector", type=float, nargs="+", default=None,
                    help="parameter vector (default: the parameter file's own values)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--timeout", type=float, default=300.0,
                    help="seconds a rank waits for the others at an exchange")
    ap.add_argument("--json", default=None, help="rank 0 writes the result here")
    args = ap.parse_args(argv)

    import torch.distributed as dist
    dist.init_process_group("gloo", timeout=datetime.timedelta(seconds=args.timeout))
    try:
        L = sep.load_library()
        sep.require_device()
        sep.set_device(int(os.environ.get("LOCAL_RANK", "0")) % L.mw_device_count())
        sep.set_precision(sep.PREC_MIXED if args.precision == "mixed" else sep.PREC_FP64)
        with ShardedEvaluator(args.params, args.stars) as ev:
            if args.vector is None:
                vec = np.zeros(ev.n_params)
                L.mw_get_search_parameters(C.byref(ev.ap), vec.ctypes.data_as(_dp))
            else:
                vec = np.asarray(args.vector, dtype=np.float64)
            lnl = ev.evaluate(vec)
            integrals = ev.last_integrals
            out = {"rank": ev.rank, "world": ev.world, "likelihood": lnl,
                   "background_integral": float(integrals[0]),
                   "stream_integrals": [float(x) for x in integrals[1:]],
                   "stars": ev.sp.number_stars, "total_stars": ev.total_stars}
        print(json.dumps(out), flush=True)
        if args.json and dist.get_rank() == 0:
            with open(args.json, "w") as f:
                json.dump(out, f)
    finally:
        dist.destroy_process_group()
    return 0


if __name__ == "__main__":
    sys.exit(main())
