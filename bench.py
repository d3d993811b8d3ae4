#!/usr/bin/env python3
"""Benchmark of the MI355X separation likelihood evaluation.

Headline (default ``--workload batch``): batched mixed-precision likelihood
evaluations against plain FP64, at per-star |d ln P| <= 1e-7.  A step is one
batch of 64 parameter vectors (4 streams each, 16 per wedge) over 4 resident
wedges of uneven size -- 70x160x64 (n=30, 10k stars), 350x400x160 (n=60, 50k
stars), 700x800x320 (n=120, 100k stars) and 700x1600x640 (n=120, 200k stars),
r x mu x nu -- evaluated in one persistent launch replayed from a HIP graph,
kernel variant chosen by the autotuner.  ``value`` is evaluations per second
of that path with the resampled quadrature (MW_PREC_RESAMPLED); its accuracy
against the FP64 path is measured in the same run and reported.  Beside it:
the same batch with the exact quadrature (MW_PREC_MIXED), the same 64 vectors
one at a time in plain FP64, the full-size FP64 evaluation (700x1600x640 plus
a 700x400x160 cut, n=120, 3 streams, 200k stars), a short Poisson serving run,
and the CPU baseline.

``--workload full`` times single full-size evaluations instead (``--precision``
fp64, mixed or resampled).  Inputs (tables, stars) are resident in HBM before every
timed region; each step uploads its parameter vectors and reads back its
results.  Prints one JSON line (the last line of stdout).  --gpus above 1
exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from milkyway_client_amd import autotune, serving, synth, workmodel
from milkyway_client_amd import separation as sep
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (what the mixed path is priced against)
BATCH_STREAMS = 4
PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def measured_traffic(key: str):
    """HBM bytes per launch of a kernel/workload, as measured by the
    rocprofv3 FETCH_SIZE + WRITE_SIZE runs summarised under profiles/ (a
    'traffic_bytes_per_launch: <key> <bytes>' line), or None."""
    try:
        for name in sorted(os.listdir(PROFILES)):
            with open(os.path.join(PROFILES, name)) as f:
                for line in f:
                    parts = line.split()
                    if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == key:
                        return int(parts[2])
    except OSError:
        pass
    return None
BATCH_VECTORS_PER_WEDGE = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --------------------------------------------------------------------------
# CPU baseline: the oracle on one core over a bounded sample, scaled
# --------------------------------------------------------------------------

def cpu_seconds_per_eval(p, stars, rows_per_region, n_star_sample):
    """Seconds of one full evaluation of (p, stars) by the oracle on one
    core, from a sample of rows_per_region mu-rows per region and
    n_star_sample stars (cost per mu-row and per star is uniform)."""
    from oracle import oracle as O
    O.build()
    ph = ParamsHolder(p)
    sample = stars[:n_star_sample]
    sh = StarsHolder(sample)
    st = StateHolder(ph, sh)
    for g in range(ph.ap.number_integrals):
        a = st.area(g)
        a.max_calculation = min(rows_per_region, a.mu_steps) * a.nu_steps * a.r_steps
    t0 = time.perf_counter()
    O.integrals(ph, sh, st)
    t_int = time.perf_counter() - t0
    sampled = sum(min(rows_per_region, r.mu_steps) * r.nu_steps * r.r_steps for r in p.regions)
    total = sum(r.mu_steps * r.nu_steps * r.r_steps for r in p.regions)
    st.es.background_integral = 1.0
    for s in range(p.number_streams):
        st.es.stream_integrals[s] = 1.0
    t0 = time.perf_counter()
    O.likelihood(ph, sh, st)
    t_star = time.perf_counter() - t0
    return t_int * total / sampled + t_star * len(stars) / len(sample), t_int + t_star


# --------------------------------------------------------------------------
# Workloads
# --------------------------------------------------------------------------

def mixed_roofline(points: float, rec: float, n_streams: int, ms: float, bg: float = None,
                   terms: float = None) -> dict:
    """Roofline of a mixed-precision integral launch, each evaluation priced
    by the method that ran it: the frozen per-point model (FP32 exp, against
    the FP32 vector peak) for exponentials per point; on distance grids the
    same weights over the recurrence's operations (FP64, against the FP64
    peak): 4 per stream term (terms: the library's count of them, n_streams
    per node on a shared grid, or the sum over streams with grids of their
    own) and 35 per background evaluation (bg: every 1st, 2nd or 4th node,
    or the background's own grid).  frac is the time the launch would take
    at those peaks over its measured time."""
    bg_rec = rec if bg is None else bg - (points - rec)
    if terms is None:
        terms = rec * n_streams
    w_exp = (points - rec) * workmodel.flops_per_point(n_streams)
    w_rec = terms * workmodel.STREAM_RECURRENCE_OPS + bg_rec * workmodel.flops_per_point(0)
    t_peak = (w_exp / FP32_PEAK_TFLOPS + w_rec / FP64_PEAK_TFLOPS) / 1e12
    work = w_exp + w_rec
    return {"achieved": round(work / (ms * 1e-3) / 1e12, 3),
            "peak": round(work / t_peak / 1e12, 2) if t_peak > 0 else FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(t_peak / (ms * 1e-3), 4),
            "integrand_evaluations_per_launch": int(points),
            "of_which_by_recurrence": int(rec),
            "stream_terms_by_recurrence": int(terms),
            "background_evaluations": int(bg_rec + points - rec),
            "algorithmic_flop_eq_per_launch": work,
            "pricing": (f"{workmodel.flops_per_point(n_streams)} flop-eq per evaluation with an exp "
                        f"per stream at {FP32_PEAK_TFLOPS} TFLOP/s (FP32); by recurrence "
                        f"{workmodel.STREAM_RECURRENCE_OPS} per stream term and "
                        f"{workmodel.flops_per_point(0)} per background evaluation at "
                        f"{FP64_PEAK_TFLOPS} (FP64); peak is their work-weighted mix")}


def build_batch():
    groups, vecs, plist = [], [], []
    for k, shape in enumerate(synth.BATCH_SHAPES):
        p, stars = synth.workload(100 + k, shape, BATCH_STREAMS)
        ctx = sep.Context(ParamsHolder(p), StarsHolder(stars))
        desc = (f"{shape.r_steps}x{shape.mu_steps}x{shape.nu_steps}n{shape.convolve}"
                f"s{BATCH_STREAMS}N{shape.n_stars}")
        groups.append((ctx, BATCH_VECTORS_PER_WEDGE, desc))
        vecs.append(synth.parameter_vectors(100 + k, p, BATCH_VECTORS_PER_WEDGE))
        plist.append((p, stars))
    return groups, vecs, plist


def full_size_fp64(steps: int, warmup: int, precision=sep.PREC_FP64):
    p, stars = synth.workload(1, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    t0 = time.perf_counter()
    ctx = sep.Context(ParamsHolder(p), StarsHolder(stars), timing=True)
    setup = time.perf_counter() - t0
    vecs = synth.parameter_vectors(1, p, 4)
    for i in range(warmup):
        ctx.eval(vecs[i % 4], precision=precision)
    lat, kern, star, points, rec, bgp = [], [], [], [], [], []
    t0 = time.perf_counter()
    for i in range(steps):
        s0 = time.perf_counter()
        ctx.eval(vecs[i % 4], precision=precision)
        lat.append(time.perf_counter() - s0)
        tm = ctx.timing()
        kern.append(tm.integral_ms)
        star.append(tm.star_ms)
        points.append(tm.integral_points)
        rec.append(tm.integral_points_recurrence)
        bgp.append(tm.integral_points_background)
    el = time.perf_counter() - t0
    ctx.close()
    # Integrand evaluations of the integral pass: cells x n for the exact
    # quadrature (workmodel.integral_points), grid nodes when resampled.
    flops = float(np.mean(points)) * workmodel.flops_per_point(synth.FULL_SIZE_STREAMS)
    k_ms = float(np.mean(kern))
    return p, stars, {
        "value": steps / el, "unit": "evals/s", "ms_per_eval": 1e3 * el / steps,
        "integral_kernel_ms": round(k_ms, 3), "star_kernel_ms": round(float(np.mean(star)), 4),
        "achieved_tflops": round(flops / (k_ms * 1e-3) / 1e12, 3),
        "frac_fp64_peak": round(flops / (k_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
        "integrand_evaluations": int(np.mean(points)), "flop_eq": flops,
        "integrand_evaluations_by_recurrence": int(np.mean(rec)),
        "background_evaluations": int(np.mean(bgp)),
        "resident_setup_s": round(setup, 3),
        "workload": "700x1600x640 + 700x400x160 cut (r x mu x nu), n=120, 3 streams, 200k stars",
    }


def serving_run(n_requests: int, rates, variant: int, bpc: int):
    """Poisson arrival streams at each rate over 8 resident wedges; requests
    are resampled mixed precision with a 10% FP64 share.  The plain
    one-at-a-time FP64 server is run at the lowest rate only (it saturates
    at about 10 requests/s)."""
    wedges = serving.make_wedges(8, seed=4000)
    holders = [(ParamsHolder(p), StarsHolder(s)) for p, s, _ in wedges]
    out = {"wedges": "8 resident wedges from the batch shapes, 1-4 streams, 10% FP64 requests",
           "sweep": []}
    with sep.Server(max_inflight=3, max_batch=32, variant=variant, blocks_per_cu=bpc) as srv:
        ids = [srv.add_wedge(ph, sh) for ph, sh in holders]
        for k, rate in enumerate(rates):
            n = max(n_requests, min(400, int(2 * rate)))
            reqs = serving.make_requests(wedges, n, seed=5000 + k, fast_precision=sep.PREC_RESAMPLED)
            arr = serving.poisson_arrivals(len(reqs), rate, seed=6000 + k)
            b0 = srv.stats().batches
            v = serving.run(srv, ids, reqs, arr)
            st = srv.stats()
            row = {"poisson_rate_per_s": rate, "requests": n,
                   "throughput_req_per_s": round(v["throughput"], 2),
                   "p50_ms": round(v["p50_ms"], 2), "p99_ms": round(v["p99_ms"], 2),
                   "batches": int(st.batches - b0)}
            for prec, name in ((sep.PREC_RESAMPLED, "resampled"), (sep.PREC_FP64, "fp64")):
                if f"p50_ms_prec{prec}" in v:
                    row[f"{name}_p50_ms"] = round(v[f"p50_ms_prec{prec}"], 2)
                    row[f"{name}_p99_ms"] = round(v[f"p99_ms_prec{prec}"], 2)
                    row[f"{name}_requests"] = v[f"n_prec{prec}"]
                    row[f"{name}_mean_queue_ms"] = round(v[f"queue_ms_prec{prec}"], 2)
                    row[f"{name}_mean_service_ms"] = round(v[f"service_ms_prec{prec}"], 2)
            log(f"  rate {rate}/s: {row}")
            out["sweep"].append(row)
        out["largest_batch"] = int(srv.stats().largest_batch)
    rate = rates[0]
    reqs = serving.make_requests(wedges, n_requests, seed=5000, fast_precision=sep.PREC_RESAMPLED)
    arr = serving.poisson_arrivals(len(reqs), rate, seed=6000)
    with sep.Server(max_inflight=1, max_batch=1) as srv:
        ids = [srv.add_wedge(ph, sh) for ph, sh in holders]
        plain = serving.run(srv, ids, reqs, arr, force_precision=sep.PREC_FP64)
    out["one_at_a_time_fp64"] = {"poisson_rate_per_s": rate, "requests": n_requests,
                                 "throughput_req_per_s": round(plain["throughput"], 2),
                                 "p50_ms": round(plain["p50_ms"], 2), "p99_ms": round(plain["p99_ms"], 2)}
    return out


def timed_batch(pairs, vecs, precision, choice, steps, warmup):
    """Eager run (kernel events around the persistent launch), then `steps`
    graph replays after `warmup`; returns evals/s, per-batch latencies,
    elapsed seconds and the eager timing."""
    n_vec = sum(n for _, n in pairs)
    with sep.Batch(pairs, precision=precision, graph=False) as eager:
        eager.set_variant(choice["variant"], choice["blocks_per_cu"])
        eager.run(vecs)
        eager.run(vecs)
        tm = eager.timing()
    with sep.Batch(pairs, precision=precision, graph=True) as b:
        b.set_variant(choice["variant"], choice["blocks_per_cu"])
        for _ in range(warmup):
            b.run(vecs)
        sep.synchronize()
        lat = []
        t0 = time.perf_counter()
        for i in range(steps):
            s0 = time.perf_counter()
            b.run(vecs)
            lat.append(time.perf_counter() - s0)
            log(f"  step {i}: batch of {n_vec} in {lat[-1] * 1e3:.1f} ms")
        sep.synchronize()
        elapsed = time.perf_counter() - t0
    return steps * n_vec / elapsed, lat, elapsed, tm


def accuracy_vs_fp64(groups, vecs, precision, per_wedge=2):
    """Largest per-star |d ln P| and integral relative errors of `precision`
    against the GPU FP64 path (itself within 1e-12 of the FP64 oracle), over
    the first vectors of every wedge."""
    worst = {"per_star_dlnP": 0.0, "integral_rel": 0.0, "dlnL": 0.0}
    for (ctx, _, _), v in zip(groups, vecs):
        rows = np.asarray(v[:per_wedge])
        ref = ctx.eval(rows, precision=sep.PREC_FP64, per_star=True)
        got = ctx.eval(rows, precision=precision, per_star=True)
        for a, b in zip(got, ref):
            worst["per_star_dlnP"] = max(worst["per_star_dlnP"],
                                         float(np.log(10) * np.max(np.abs(a.log10_p - b.log10_p))))
            ints_a = np.concatenate([[a.background_integral], a.stream_integrals])
            ints_b = np.concatenate([[b.background_integral], b.stream_integrals])
            worst["integral_rel"] = max(worst["integral_rel"],
                                        float(np.max(np.abs(ints_a - ints_b) / np.abs(ints_b))))
            worst["dlnL"] = max(worst["dlnL"], float(np.log(10) * abs(a.likelihood - b.likelihood)))
    return {k: float(f"{v:.3e}") for k, v in worst.items()}


def bench_batch(args):
    t0 = time.perf_counter()
    groups, vecs, plist = build_batch()
    log(f"batch workload resident in {time.perf_counter() - t0:.2f} s")
    pairs = [(c, n) for c, n, _ in groups]
    n_vec = sum(n for _, n in pairs)

    head_prec = sep.PREC_RESAMPLED if args.batch_precision == "resampled" else sep.PREC_MIXED
    other_prec = sep.PREC_MIXED if head_prec == sep.PREC_RESAMPLED else sep.PREC_RESAMPLED
    tune_t0 = time.perf_counter()
    if args.variant is not None:
        choice = {"variant": args.variant[0], "blocks_per_cu": args.variant[1], "table": []}
    else:
        choice = autotune.tune(groups, vecs, head_prec, proxy_vectors=1, reps=1,
                               blocks_per_cu=(2, 4, 5), use_cache=False, log=log)
    tune_s = time.perf_counter() - tune_t0
    log(f"autotuner ({args.batch_precision}): variant {choice['variant']} blocks/CU "
        f"{choice['blocks_per_cu']} ({tune_s:.1f} s)")
    log(f"batched mixed precision, headline ({args.batch_precision})")
    value, lat, elapsed, tm = timed_batch(pairs, vecs, head_prec, choice, args.steps, args.warmup)
    int_ms, points = tm.integral_ms, tm.integral_points
    roof = mixed_roofline(float(points), float(tm.integral_points_recurrence), BATCH_STREAMS, int_ms,
                          float(tm.integral_points_background), float(tm.integral_stream_terms))
    log(f"  eager: persistent integral launch {int_ms:.2f} ms of {tm.total_ms:.2f} ms, "
        f"{points:.3e} integrand evaluations")

    second = None
    if not args.no_secondary:
        tune2 = autotune.tune(groups, vecs, other_prec, proxy_vectors=1, reps=1,
                              blocks_per_cu=(2, 4, 5), use_cache=False, log=log)
        log("batched mixed precision, the other quadrature")
        v2, lat2, _, tm2 = timed_batch(pairs, vecs, other_prec, tune2, max(1, min(args.steps, 2)), 1)
        second = (v2, lat2, tm2, tune2)

    # Plain FP64, one vector at a time, same vectors.
    log("plain FP64, one at a time")
    t0 = time.perf_counter()
    for (ctx, n, _), v in zip(groups, vecs):
        for row in v:
            ctx.eval(row, precision=sep.PREC_FP64)
    sep.synchronize()
    plain = n_vec / (time.perf_counter() - t0)
    log(f"  {plain:.3f} evals/s")

    log("accuracy against FP64")
    acc = accuracy_vs_fp64(groups, vecs, head_prec)
    log(f"  {args.batch_precision}: {acc}")

    quad = "resampled" if head_prec == sep.PREC_RESAMPLED else "exact"
    quad_desc = ("resampled: each region's r-step x convolution weights folded by 16-node Lagrange "
                 "interpolation onto grids sized to the model's features (MW_PREC_RESAMPLED): each "
                 "stream's Gaussian by FP64 recurrence on a distance grid of its own (rows it does "
                 "not reach skipped), the background on a magnitude grid of its own; star pass: "
                 "streams at every convolution point, background folded over the convolution "
                 "offsets" if quad == "resampled" else
                 "exact: the reference's r-step x convolution points (MW_PREC_MIXED)")
    out = {
        "metric": "likelihood evals/s, batched mixed precision vs plain FP64 @ per-star |d lnP| <= 1e-7",
        "value": round(value, 4), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (seeded wedge geometries, parameter vectors and stars)",
        "config": {
            "workload": ("batch: 64 parameter vectors (4 streams, 16 per wedge) over 4 wedges "
                         "70x160x64 n=30 10k stars, 350x400x160 n=60 50k, 700x800x320 n=120 100k, "
                         "700x1600x640 n=120 200k (r x mu x nu); one persistent launch "
                         "replayed from a HIP graph"),
            "batch": n_vec, "precision": "mixed",
            "quadrature": quad_desc,
            "kernel_variant": choice["variant"], "blocks_per_cu": choice["blocks_per_cu"],
            "parallelism": "single GPU",
        },
        "accuracy_vs_fp64": dict(acc, bar_per_star_dlnP=1e-7,
                                 what="worst over 2 vectors per wedge, against the GPU FP64 path"),
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 2),
                       "what": "one batch of 64 evaluations"},
        "plain_fp64": {"value": round(plain, 4), "unit": "evals/s",
                       "ratio_batched_mixed_over_plain_fp64": round(value / plain, 3),
                       "what": "the same 64 vectors, one evaluation at a time, FP64, exact quadrature"},
        "roofline": dict(
            {"bound": "valu", "kernel": f"integral_persistent (mixed, {quad} quadrature)"},
            **roof,
            traffic=measured_traffic(f"integral_persistent_{args.batch_precision}_batch64"),
            avg_launch_ms=round(int_ms, 3)),
        "tuning": {"seconds": round(tune_s, 2), "table": choice["table"]},
    }
    if second:
        v2, lat2, tm2, tune2 = second
        q2 = "exact" if quad == "resampled" else "resampled"
        roof2 = mixed_roofline(float(tm2.integral_points), float(tm2.integral_points_recurrence),
                               BATCH_STREAMS, tm2.integral_ms, float(tm2.integral_points_background),
                               float(tm2.integral_stream_terms))
        key2 = "mixed" if q2 == "exact" else "resampled"
        out[f"batched_mixed_{q2}_quadrature"] = {
            "value": round(v2, 4), "unit": "evals/s", "ratio_over_plain_fp64": round(v2 / plain, 3),
            "accuracy_vs_fp64": accuracy_vs_fp64(groups, vecs, other_prec, per_wedge=1),
            "p50_ms": round(1e3 * float(np.percentile(lat2, 50)), 2),
            "kernel_variant": tune2["variant"], "blocks_per_cu": tune2["blocks_per_cu"],
            "roofline": dict({"bound": "valu", "kernel": f"integral_persistent (mixed, {q2} quadrature)"},
                             **roof2, traffic=measured_traffic(f"integral_persistent_{key2}_batch64"),
                             avg_launch_ms=round(tm2.integral_ms, 3)),
        }
    for ctx, *_ in groups:
        ctx.close()

    if not args.no_full:
        log("full-size FP64 evaluation")
        _, _, full = full_size_fp64(3, 1)
        out["full_size_fp64"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in full.items()}
    if not args.no_serving:
        log("serving run")
        out["serving"] = serving_run(args.serving_requests, args.serving_rates, choice["variant"],
                                     choice["blocks_per_cu"])
    if not args.no_cpu:
        log("CPU baseline (oracle, 1 thread, bounded sample)")
        total_s, spent = 0.0, 0.0
        for p, stars in plist:
            s, used = cpu_seconds_per_eval(p, stars, 6, 2000)
            total_s += s * BATCH_VECTORS_PER_WEDGE
            spent += used
        out["cpu_baseline"] = {
            "value": round(n_vec / total_s, 8), "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": (f"oracle FP64 (C restatement of the reference CPU path), 1 thread: six mu-rows of "
                       f"every region of each of the 4 wedges and 2000 stars each ({spent:.1f} s), "
                       f"scaled linearly to the 64-vector batch = {total_s:.0f} s"),
        }
    else:
        out["cpu_baseline"] = None
    return out


def bench_full(args):
    prec = {"fp64": sep.PREC_FP64, "mixed": sep.PREC_MIXED, "resampled": sep.PREC_RESAMPLED}[args.precision]
    p, stars, full = full_size_fp64(args.steps, args.warmup, prec)
    flops = full["flop_eq"]
    out = {
        "metric": "likelihood evals/s (separation integral + star pass)",
        "value": round(full["value"], 4), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(full["ms_per_eval"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == sep.PREC_FP64 else "f64+f32",
        "data": "synthetic (seeded wedge geometry, parameters and stars)",
        "integrand_evaluations": full["integrand_evaluations"],
        "config": {"workload": "full: " + full["workload"], "precision": args.precision,
                   "batch": 1, "parallelism": "single GPU"},
        "star_kernel_ms": full["star_kernel_ms"],
        "resident_setup_s": full["resident_setup_s"],
        "roofline": ({"bound": "valu", "kernel": "integral_kernel", "achieved": full["achieved_tflops"],
                      "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": round(full["achieved_tflops"] / FP64_PEAK_TFLOPS, 4),
                      "traffic": measured_traffic("integral_kernel_fp64_full"),
                      "algorithmic_flop_eq_per_launch": flops,
                      "avg_launch_ms": full["integral_kernel_ms"]}
                     if prec == sep.PREC_FP64 else
                     dict({"bound": "valu", "kernel": "integral_kernel"},
                          **mixed_roofline(float(full["integrand_evaluations"]),
                                           float(full["integrand_evaluations_by_recurrence"]),
                                           synth.FULL_SIZE_STREAMS, full["integral_kernel_ms"],
                                           float(full["background_evaluations"])),
                          traffic=None, avg_launch_ms=full["integral_kernel_ms"])),
    }
    if not args.no_cpu:
        s, spent = cpu_seconds_per_eval(p, stars, 4, 4000)
        out["cpu_baseline"] = {
            "value": round(1.0 / s, 8), "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": (f"oracle FP64, 1 thread: 4 mu-rows of every region and 4000 stars ({spent:.1f} s), "
                       f"scaled linearly to one full evaluation = {s:.0f} s")}
    else:
        out["cpu_baseline"] = None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["batch", "full"], default="batch")
    ap.add_argument("--precision", choices=["fp64", "mixed", "resampled"], default="fp64",
                    help="precision of --workload full")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-full", action="store_true", help="skip the full-size FP64 line")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving run")
    ap.add_argument("--batch-precision", choices=["resampled", "mixed"], default="resampled",
                    help="precision of the headline batch (the other one runs beside it)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the second batch precision (for profiling one path)")
    ap.add_argument("--variant", type=int, nargs=2, metavar=("V", "BLOCKS_PER_CU"), default=None,
                    help="pin the headline batch's kernel variant and blocks per CU instead of "
                         "tuning (counter runs of one kernel: the tuner's timings move under "
                         "a profiler)")
    ap.add_argument("--serving-requests", type=int, default=40)
    ap.add_argument("--serving-rates", type=float, nargs="+", default=[10.0, 50.0, 200.0, 800.0],
                    help="Poisson arrival rates (requests/s) of the serving sweep")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build evaluates on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    sep.load_library()
    sep.require_device()
    out = bench_batch(args) if args.workload == "batch" else bench_full(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
