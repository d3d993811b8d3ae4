"""Serving workload: a Poisson stream of likelihood requests over resident
wedges, driven against the request server (mw_server_*).

Requests are submitted at their arrival times by the calling thread (the
server never blocks a submit); latency is completion minus submission on the
server's clock.  The comparison point is the same arrival stream served one
request at a time in plain FP64 (a server with one runner and batches of one).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

from . import separation as sep
from . import synth
from .refstructs import ParamsHolder, StarsHolder


@dataclass
class Request:
    wedge: int
    vector: np.ndarray
    precision: int


def make_wedges(n_wedges: int = 8, seed: int = 4000, shapes=None, max_streams: int = 4):
    """n_wedges resident-wedge descriptions drawn from the batch shapes, with
    1..max_streams streams each (seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shapes = shapes or synth.BATCH_SHAPES
    out = []
    for i in range(n_wedges):
        shape = shapes[i % len(shapes)]
        S = int(rng.integers(1, max_streams + 1))
        p, stars = synth.workload(seed + i, shape, S)
        out.append((p, stars, shape))
    return out


def make_requests(wedges, n: int, seed: int, fp64_fraction: float = 0.1,
                  fast_precision: int = sep.PREC_MIXED) -> List[Request]:
    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    for j in range(n):
        w = int(rng.integers(0, len(wedges)))
        p = wedges[w][0]
        vec = synth.parameter_vectors(seed + j, p, 1)[0]
        prec = sep.PREC_FP64 if rng.random() < fp64_fraction else fast_precision
        reqs.append(Request(w, vec, prec))
    return reqs


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, n))


def run(server: sep.Server, wedge_ids: Sequence[int], requests: Sequence[Request],
        arrivals: Sequence[float], force_precision: int = None) -> dict:
    """Submits requests[i] at arrivals[i] seconds after the start; waits for
    all; returns latencies and throughput."""
    t0 = time.perf_counter()
    ids = []
    for req, at in zip(requests, arrivals):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        prec = req.precision if force_precision is None else force_precision
        ids.append(server.submit(wedge_ids[req.wedge], req.vector, prec))
    lat, first, last, results, precs, queued = [], None, None, [], [], []
    for req, rid in zip(requests, ids):
        res, (ts, td, tdone) = server.wait(rid)
        lat.append(tdone - ts)
        queued.append(td - ts)
        first = ts if first is None else min(first, ts)
        last = tdone if last is None else max(last, tdone)
        results.append(res)
        precs.append(req.precision if force_precision is None else force_precision)
    lat = np.asarray(lat)
    queued = np.asarray(queued)
    precs = np.asarray(precs)
    out = {
        "n": len(ids),
        "throughput": len(ids) / (last - first) if last > first else float("nan"),
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "mean_ms": 1e3 * float(lat.mean()),
        "results": results,
    }
    # the same percentiles per precision class
    for prec in sorted(set(precs.tolist())):
        sel = lat[precs == prec]
        out[f"p50_ms_prec{prec}"] = 1e3 * float(np.percentile(sel, 50))
        out[f"p99_ms_prec{prec}"] = 1e3 * float(np.percentile(sel, 99))
        out[f"n_prec{prec}"] = int(sel.size)
        # split: waiting for admission vs dispatched until collected
        out[f"queue_ms_prec{prec}"] = 1e3 * float(queued[precs == prec].mean())
        out[f"service_ms_prec{prec}"] = 1e3 * float((sel - queued[precs == prec]).mean())
    return out
