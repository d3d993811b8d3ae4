#!/usr/bin/env python3
"""Benchmark of the MI355X separation likelihood evaluation.

Headline (default ``--workload batch``): batched mixed-precision likelihood
evaluations on the reference's own quadrature, against plain FP64, at
per-star |d ln P| <= 1e-7.  A step is one batch of 64 parameter vectors
(4 streams each, 16 per wedge) over 4 resident wedges of uneven size --
70x160x64 (n=30, 10k stars), 350x400x160 (n=60, 50k stars), 700x800x320
(n=120, 100k stars) and 700x1600x640 (n=120, 200k stars), r x mu x nu --
evaluated in one persistent launch replayed from a HIP graph, kernel variant
chosen by the autotuner.  Precision MW_PREC_MIXED: every r-step x
convolution point of every cell is evaluated, as the reference does.  Every
timed step evaluates a fresh set of 64 vectors.  ``value`` is evaluations
per second; the accuracy of all 64 vectors of the first timed set against
the FP64 path is measured in the same run.

Beside it:
  plain_fp64      the same 64 vectors one at a time in FP64 (the 20x target's
                  denominator and the accuracy reference)
  resampled       the same batch with MW_PREC_RESAMPLED, a different
                  quadrature (the reference's weights folded onto coarser
                  grids): fresh vectors every timed step, so re-plans, graph
                  re-captures and table folds are timed; accuracy on all 64
                  vectors of the first set; priced by the method's own
                  operation count
  full_size_fp64  one full-size FP64 evaluation (700x1600x640 plus a
                  700x400x160 cut, n=120, 3 streams, 200k stars): 3 warm-ups,
                  20 repetitions, median
  serving         a Poisson request stream over 8 resident wedges
  cpu_baseline    the oracle (C restatement of the reference CPU path) on the
                  host's cores: one core and all cores, on a sample of the
                  headline batch, on configs 70x160x64 in full and on a slice
                  of the full-size wedge

``--workload full`` times single full-size evaluations instead (``--precision``
fp64, mixed or resampled).  Inputs (tables, stars) are resident in HBM before
every timed region; each step uploads its parameter vectors and reads back
its results.  Prints one JSON line (the last line of stdout).  --gpus above 1
exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import subprocess
import sys
import time

import numpy as np

from milkyway_client_amd import autotune, serving, synth, workmodel
from milkyway_client_amd import separation as sep
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (what the mixed path is priced against)
BATCH_STREAMS = 4
BATCH_VECTORS_PER_WEDGE = 16
LN10 = float(np.log(10.0))
PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")

# The resampled method's own operation count (not the frozen model, whose
# integrand evaluations it replaces by fewer grid nodes): per stream term on
# a distance grid the multiply-add and the two recurrence multiplies, no
# exponential; per background evaluation the frozen model's 35.
RESAMPLED_STREAM_TERM_OPS = 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _profile_lines(tag: str):
    try:
        for name in sorted(os.listdir(PROFILES)):
            with open(os.path.join(PROFILES, name)) as f:
                for line in f:
                    parts = line.split()
                    if parts and parts[0] == tag:
                        yield name, parts[1:]
    except OSError:
        return


def measured_traffic(key: str):
    """HBM bytes per launch of a kernel/workload, as measured by the
    rocprofv3 FETCH_SIZE + WRITE_SIZE runs summarised under profiles/ (a
    'traffic_bytes_per_launch: <key> <bytes>' line), or None."""
    for _, parts in _profile_lines("traffic_bytes_per_launch:"):
        if len(parts) == 2 and parts[0] == key:
            return int(parts[1])
    return None


def measured_counters(key: str):
    """VALU counter figures of a kernel/workload from the profile summaries
    (a 'valu_counters: <key> name=value ...' line), or None."""
    for name, parts in _profile_lines("valu_counters:"):
        if parts and parts[0] == key:
            out = {}
            for kv in parts[1:]:
                k, _, v = kv.partition("=")
                out[k] = float(v)
            out["source"] = f"profiles/{name}"
            return out
    return None


# --------------------------------------------------------------------------
# CPU baseline: the oracle on the host's cores over bounded samples, scaled
# --------------------------------------------------------------------------

def cpu_description(threads: int) -> dict:
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        cc = "gcc"
    return {"cpu": model, "logical_cpus": os.cpu_count(), "threads_all_cores": threads,
            "compiler": cc,
            "flags": "-O3 -std=gnu11 -ffp-contract=off -fno-fast-math -fopenmp (oracle/Makefile)"}


def first_rows(p, rows: int):
    """Copy of p whose regions keep only their first `rows` mu rows (same
    cells, same work per cell)."""
    q = copy.deepcopy(p)
    for g in q.regions:
        k = min(rows, g.mu_steps)
        step = (g.mu_max - g.mu_min) / g.mu_steps
        g.mu_max = g.mu_min + k * step
        g.mu_steps = k
    return q


def rows_share(p, rank: int, ranks: int):
    """Copy of p whose regions keep rank's contiguous share of their mu rows
    (rows [rank * m / ranks, (rank + 1) * m / ranks) of each region's m, same
    cells): one GPU's part of a row-aligned split over `ranks` GPUs, as the
    flat calculation-index split of parameters.c:502-522 gives it."""
    q = copy.deepcopy(p)
    for g in q.regions:
        step = (g.mu_max - g.mu_min) / g.mu_steps
        k0, k1 = rank * g.mu_steps // ranks, (rank + 1) * g.mu_steps // ranks
        g.mu_min, g.mu_max, g.mu_steps = g.mu_min + k0 * step, g.mu_min + k1 * step, k1 - k0
    return q


def cpu_seconds_per_eval(p, stars, rows: int, n_star_sample: int, threads: int):
    """Seconds of one full evaluation of (p, stars) by the oracle (threads 0:
    the serial restatement of the reference; else its OpenMP mode over mu
    rows), from the first `rows` mu rows of every region and n_star_sample
    stars; cost per mu row and per star is uniform.  Returns (scaled
    seconds, seconds spent)."""
    from oracle import oracle as O
    O.build()
    sample_p = first_rows(p, rows)
    ph = ParamsHolder(sample_p)
    sh = StarsHolder(stars[:n_star_sample])
    st = StateHolder(ph, sh)
    t0 = time.perf_counter()
    O.integrals(ph, sh, st, threads=threads)
    t_int = time.perf_counter() - t0
    sampled = sum(g.r_steps * g.mu_steps * g.nu_steps for g in sample_p.regions)
    total = sum(g.r_steps * g.mu_steps * g.nu_steps for g in p.regions)
    st.es.background_integral = 1.0
    for s in range(p.number_streams):
        st.es.stream_integrals[s] = 1.0
    t0 = time.perf_counter()
    O.likelihood(ph, sh, st, threads=threads)
    t_star = time.perf_counter() - t0
    n_sample = min(n_star_sample, len(stars))
    return t_int * total / sampled + t_star * len(stars) / n_sample, t_int + t_star


def cpu_baseline(plist, threads: int) -> dict:
    """The oracle on one core and on `threads` cores: the headline batch
    (sampled, scaled to 64 vectors), the 70x160x64 workload (n = 30, 1
    stream, 10k stars) in full, and a slice of the full-size wedge (16 of
    1600 main mu rows plus 4 of 400 cut rows, 4000 stars) scaled to one
    evaluation."""
    out = cpu_description(threads)
    n_vec = BATCH_VECTORS_PER_WEDGE * len(plist)
    modes = (("one_core", 0, 6, 2000), ("all_cores", threads, max(6, 3 * threads), 20000))
    batch = {}
    for name, th, rows, ns in modes:
        total, spent = 0.0, 0.0
        for p, stars in plist:
            s, used = cpu_seconds_per_eval(p, stars, rows, ns, th)
            total += s * BATCH_VECTORS_PER_WEDGE
            spent += used
        batch[name] = {"value": round(n_vec / total, 8), "unit": "evals/s", "cores": max(th, 1),
                       "sample": (f"{rows} mu-rows of every region of each of the 4 wedges and "
                                  f"{ns} stars each ({spent:.1f} s), scaled linearly to the "
                                  f"64-vector batch = {total:.0f} s")}
        log(f"  CPU batch sample, {name}: {batch[name]}")
    p0, s0 = synth.workload(0, synth.REFERENCE_SIZE, 1)
    cfg0 = {}
    for name, th, _, _ in modes:
        s, _ = cpu_seconds_per_eval(p0, s0, p0.regions[0].mu_steps, len(s0), th)
        cfg0[name] = {"evals_per_s": round(1.0 / s, 4), "s_per_eval": round(s, 4), "cores": max(th, 1)}
    log(f"  CPU 70x160x64 n=30 1 stream 10k stars: {cfg0}")
    p1, s1 = synth.workload(1, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    cfg1 = {}
    for name, th, _, _ in modes:
        q = first_rows(p1, 16)
        q.regions[1].mu_steps = 4
        q.regions[1].mu_max = p1.regions[1].mu_min + 4 * (p1.regions[1].mu_max - p1.regions[1].mu_min) / 400
        ph = ParamsHolder(q)
        from oracle import oracle as O
        sh = StarsHolder(s1[:4000])
        st = StateHolder(ph, sh)
        t0 = time.perf_counter()
        O.integrals(ph, sh, st, threads=th)
        t_int = time.perf_counter() - t0
        st.es.background_integral = 1.0
        for s in range(p1.number_streams):
            st.es.stream_integrals[s] = 1.0
        t0 = time.perf_counter()
        O.likelihood(ph, sh, st, threads=th)
        t_star = time.perf_counter() - t0
        scaled = t_int * 100.0 + t_star * 50.0   # 16/1600 and 4/400 of the rows; 4000 of 200k stars
        cfg1[name] = {"evals_per_s": round(1.0 / scaled, 8), "s_per_eval": round(scaled, 1),
                      "cores": max(th, 1), "sample_s": round(t_int + t_star, 1)}
    log(f"  CPU full-size slice: {cfg1}")
    out.update({
        "value": batch["all_cores"]["value"], "unit": "evals/s", "cores": threads, "kind": "port",
        "sample": ("oracle FP64 (C restatement of the reference CPU path), OpenMP over mu rows: "
                   + batch["all_cores"]["sample"]),
        "one_core": batch["one_core"],
        "reference_size_70x160x64_full": cfg0,
        "full_size_slice_16_of_1600_plus_4_of_400_rows": cfg1,
    })
    return out


# --------------------------------------------------------------------------
# Workloads
# --------------------------------------------------------------------------

def frozen_roofline(points: float, n_streams: int, ms: float) -> dict:
    """The frozen per-point model (35 + 42 S flop-eq per integrand
    evaluation) over the measured launch time, against the FP32 vector peak
    (the mixed path's exp is FP32)."""
    work = points * workmodel.flops_per_point(n_streams)
    achieved = work / (ms * 1e-3) / 1e12
    return {"achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "integrand_evaluations_per_launch": int(points),
            "algorithmic_flop_eq_per_launch": work,
            "model": (f"frozen: {workmodel.flops_per_point(n_streams)} flop-eq per integrand "
                      f"evaluation at {n_streams} streams (15 ops + sqrt 10 + divide 10, and 22 ops + "
                      f"exp 20 per stream)")}


def resampled_roofline(points: float, rec: float, n_streams: int, ms: float, bg: float,
                       terms: float) -> dict:
    """The resampled method priced by its own operation count (it evaluates
    grid nodes, not the reference's integrand points): nodes with an
    exponential per stream at the frozen per-point weights against the FP32
    peak; on distance grids RESAMPLED_STREAM_TERM_OPS per stream term and 35
    per background evaluation against the FP64 peak.  frac is the time the
    launch would take at those peaks over its measured time."""
    bg_rec = bg - (points - rec)
    w_exp = (points - rec) * workmodel.flops_per_point(n_streams)
    w_rec = terms * RESAMPLED_STREAM_TERM_OPS + bg_rec * workmodel.flops_per_point(0)
    t_peak = (w_exp / FP32_PEAK_TFLOPS + w_rec / FP64_PEAK_TFLOPS) / 1e12
    work = w_exp + w_rec
    return {"achieved": round(work / (ms * 1e-3) / 1e12, 3),
            "peak": round(work / t_peak / 1e12, 2) if t_peak > 0 else FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(t_peak / (ms * 1e-3), 4),
            "grid_nodes_per_launch": int(points),
            "of_which_by_recurrence": int(rec),
            "stream_terms_by_recurrence": int(terms),
            "background_evaluations": int(bg),
            "op_count_per_launch": work,
            "model": (f"the resampled method's own operation count, not the frozen model: "
                      f"{workmodel.flops_per_point(n_streams)} per node with an exp per stream "
                      f"(FP32 peak); by recurrence {RESAMPLED_STREAM_TERM_OPS} per stream term and "
                      f"{workmodel.flops_per_point(0)} per background evaluation (FP64 peak)")}


def vector_set(m: int, plist):
    """Parameter-vector set m of the batch: 16 seeded vectors per wedge
    (wedge k, set m: synth.parameter_vectors(100 + k + 4 m, ...))."""
    return [synth.parameter_vectors(100 + k + 4 * m, p, BATCH_VECTORS_PER_WEDGE)
            for k, (p, _) in enumerate(plist)]


def build_batch():
    groups, plist = [], []
    for k, shape in enumerate(synth.BATCH_SHAPES):
        p, stars = synth.workload(100 + k, shape, BATCH_STREAMS)
        ctx = sep.Context(ParamsHolder(p), StarsHolder(stars))
        desc = (f"{shape.r_steps}x{shape.mu_steps}x{shape.nu_steps}n{shape.convolve}"
                f"s{BATCH_STREAMS}N{shape.n_stars}")
        groups.append((ctx, BATCH_VECTORS_PER_WEDGE, desc))
        plist.append((p, stars))
    return groups, plist


def full_size_fp64(steps: int, warmup: int, precision=sep.PREC_FP64):
    p, stars = synth.workload(1, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    t0 = time.perf_counter()
    ctx = sep.Context(ParamsHolder(p), StarsHolder(stars), timing=True)
    setup = time.perf_counter() - t0
    vecs = synth.parameter_vectors(1, p, 4)
    for i in range(warmup):
        ctx.eval(vecs[i % 4], precision=precision)
    lat, kern, star, points, rec, bgp = [], [], [], [], [], []
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
    ctx.close()
    med = float(np.median(lat))
    k_ms = float(np.median(kern))
    flops = float(np.mean(points)) * workmodel.flops_per_point(synth.FULL_SIZE_STREAMS)
    # the star pass beside it: VALU-bound by the frozen model, its HBM rate
    # from the rocprofv3 traffic of the same launch where it was measured
    s_ms = float(np.median(star))
    s_flops = workmodel.star_flops(len(stars), p.convolve, synth.FULL_SIZE_STREAMS)
    s_bytes = workmodel.star_bytes(len(stars))
    s_traffic = measured_traffic("star_kernel_fp64_full") if precision == sep.PREC_FP64 else None
    star_pass = {
        "bound": "valu", "kernel": "star_kernel (FP64, one work-item per star)",
        "achieved": round(s_flops / (s_ms * 1e-3) / 1e12, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(s_flops / (s_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
        "algorithmic_bytes": s_bytes, "traffic": s_traffic,
        "hbm_GBps": round((s_traffic or s_bytes) / (s_ms * 1e-3) / 1e9, 1),
        "hbm_GBps_of": "measured traffic" if s_traffic else "algorithmic bytes",
        "hardware": measured_counters("star_kernel_fp64_full") if precision == sep.PREC_FP64 else None,
    }
    return p, stars, {
        "value": 1.0 / med, "unit": "evals/s", "ms_per_eval_median": 1e3 * med,
        "ms_per_eval_p99": 1e3 * float(np.percentile(lat, 99)),
        "warmup": warmup, "repetitions": steps,
        "integral_kernel_ms_median": round(k_ms, 3),
        "star_kernel_ms_median": round(s_ms, 4),
        "star_pass": star_pass,
        "achieved_tflops": round(flops / (k_ms * 1e-3) / 1e12, 3),
        "frac_fp64_peak": round(flops / (k_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
        "integrand_evaluations": int(np.mean(points)), "flop_eq": flops,
        "integrand_evaluations_by_recurrence": int(np.mean(rec)),
        "background_evaluations": int(np.mean(bgp)),
        "resident_setup_s": round(setup, 3),
        "workload": "700x1600x640 + 700x400x160 cut (r x mu x nu), n=120, 3 streams, 200k stars",
    }


SERVING_SET = 400


def reference_size_fp64(steps: int = 20, warmup: int = 3) -> dict:
    """One FP64 evaluation of the 70x160x64 workload (n = 30, 1 stream, 10k
    stars): median over `steps` repetitions after `warmup`."""
    p, stars = synth.workload(0, synth.REFERENCE_SIZE, 1)
    with sep.Context(ParamsHolder(p), StarsHolder(stars)) as ctx:
        v = p.parameter_vector()
        for _ in range(warmup):
            ctx.eval(v, precision=sep.PREC_FP64)
        lat = []
        for _ in range(steps):
            s0 = time.perf_counter()
            ctx.eval(v, precision=sep.PREC_FP64)
            lat.append(time.perf_counter() - s0)
    med = float(np.median(lat))
    return {"value": round(1.0 / med, 2), "unit": "evals/s", "ms_per_eval_median": round(1e3 * med, 3),
            "ms_per_eval_p99": round(1e3 * float(np.percentile(lat, 99)), 3),
            "workload": "70x160x64 (r x mu x nu), n=30, 1 stream, 10k stars", "warmup": warmup,
            "repetitions": steps}


def one_gpu_share(full_ms: float, ranks: int = 8, steps: int = 10, warmup: int = 3) -> dict:
    """One GPU's share of the full-size evaluation split over `ranks` GPUs
    (rank 0: its mu rows of both regions and 1/ranks of the stars), at its
    real shape, FP64 and mixed; the composing exchange of 1 + S doubles per
    region is not part of it (one GPU here)."""
    p, stars = synth.workload(1, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    q = rows_share(p, 0, ranks)
    sh = stars[: len(stars) // ranks]
    vecs = synth.parameter_vectors(1, p, 4)
    out = {"workload": (f"rank 0 of {ranks}: " + " + ".join(
        f"{g.r_steps}x{g.mu_steps}x{g.nu_steps}" for g in q.regions)
        + f" (r x mu x nu), n={q.convolve}, {synth.FULL_SIZE_STREAMS} streams, {len(sh)} stars")}
    with sep.Context(ParamsHolder(q), StarsHolder(sh), timing=True) as ctx:
        for name, prec in (("fp64", sep.PREC_FP64), ("mixed", sep.PREC_MIXED)):
            for i in range(warmup):
                ctx.eval(vecs[i % 4], precision=prec)
            lat, kern = [], []
            for i in range(steps):
                s0 = time.perf_counter()
                ctx.eval(vecs[i % 4], precision=prec)
                lat.append(time.perf_counter() - s0)
                kern.append(ctx.timing().integral_ms)
            out[f"{name}_ms_per_eval_median"] = round(1e3 * float(np.median(lat)), 3)
            out[f"{name}_integral_kernel_ms_median"] = round(float(np.median(kern)), 3)
    out["fp64_efficiency_vs_full_over_ranks"] = round(full_ms / ranks / out["fp64_ms_per_eval_median"], 3)
    out["what"] = ("the full-size FP64 evaluation's time / ranks over the share's: below 1 where the "
                   "share's blocks no longer fill the GPU")
    return out


def full_size_mixed_32(reps: int = 3) -> dict:
    """32 parameter vectors on the full-size wedge in one mixed-precision
    evaluation call (the exact quadrature): evaluations per second, median
    of `reps` calls after one warm-up."""
    p, stars = synth.workload(1, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    vecs = synth.parameter_vectors(2, p, 32)
    with sep.Context(ParamsHolder(p), StarsHolder(stars)) as ctx:
        ctx.eval(vecs, precision=sep.PREC_MIXED)
        lat = []
        for _ in range(reps):
            s0 = time.perf_counter()
            ctx.eval(vecs, precision=sep.PREC_MIXED)
            lat.append(time.perf_counter() - s0)
    med = float(np.median(lat))
    return {"value": round(32 / med, 3), "unit": "evals/s", "s_per_call_median": round(med, 3),
            "workload": ("32 vectors on 700x1600x640 + 700x400x160 cut, n=120, 3 streams, 200k stars, "
                         "one mw_ctx_eval call, mixed precision, exact quadrature"),
            "accuracy": "tests/test_gpu_configs.py::test_full_size_mixed_32_vectors (per star vs FP64)"}


def serving_sweep(srv, ids, all_reqs, rates, n_requests, fast):
    """Serves the first requests of all_reqs at each Poisson rate on the
    running server; one row per rate, with the fast class (named by fast) and
    the FP64 class reported apart."""
    rows = []
    for k, rate in enumerate(rates):
        n = max(n_requests, min(SERVING_SET, int(2 * rate)))
        reqs = all_reqs[:n]
        arr = serving.poisson_arrivals(len(reqs), rate, seed=6000 + k)
        b0 = srv.stats().batches
        v = serving.run(srv, ids, reqs, arr)
        st = srv.stats()
        row = {"poisson_rate_per_s": rate, "requests": n,
               "throughput_req_per_s": round(v["throughput"], 2),
               "p50_ms": round(v["p50_ms"], 2), "p99_ms": round(v["p99_ms"], 2),
               "batches": int(st.batches - b0)}
        for prec, name in (fast, (sep.PREC_FP64, "fp64")):
            if f"p50_ms_prec{prec}" in v:
                row[f"{name}_p50_ms"] = round(v[f"p50_ms_prec{prec}"], 2)
                row[f"{name}_p99_ms"] = round(v[f"p99_ms_prec{prec}"], 2)
                row[f"{name}_requests"] = v[f"n_prec{prec}"]
                row[f"{name}_mean_queue_ms"] = round(v[f"queue_ms_prec{prec}"], 2)
                row[f"{name}_mean_service_ms"] = round(v[f"service_ms_prec{prec}"], 2)
        log(f"  rate {rate}/s: {row}")
        rows.append(row)
    return rows


def serving_run(n_requests: int, rates, exact_rates, variant: int, bpc: int):
    """Poisson arrival streams at each rate over 8 resident wedges, 10% of
    the requests FP64.  Two sweeps over one seeded request sequence (the same
    wedges, vectors and FP64 flags): the other requests in resampled mixed
    precision (`sweep`), then in mixed precision on the exact quadrature
    (`exact_quadrature_sweep`, the reference's r-step x convolution points of
    every cell).  Every rate serves the sequence's first
    max(n_requests, min(400, 2 rate)) requests, so that only the arrival
    rate changes between rows.  The plain one-at-a-time FP64 server is run
    at the lowest rate only (it saturates at about 10 requests/s)."""
    wedges = serving.make_wedges(8, seed=4000)
    holders = [(ParamsHolder(p), StarsHolder(s)) for p, s, _ in wedges]
    out = {"wedges": "8 resident wedges from the batch shapes, 1-4 streams, 10% FP64 requests",
           "requests": "one seeded sequence of 400 requests; each rate serves its first n"}
    resampled = serving.make_requests(wedges, SERVING_SET, seed=5000, fast_precision=sep.PREC_RESAMPLED)
    exact = serving.make_requests(wedges, SERVING_SET, seed=5000, fast_precision=sep.PREC_MIXED)
    with sep.Server(max_inflight=3, max_batch=32, variant=variant, blocks_per_cu=bpc) as srv:
        ids = [srv.add_wedge(ph, sh) for ph, sh in holders]
        out["sweep"] = serving_sweep(srv, ids, resampled, rates, n_requests,
                                     (sep.PREC_RESAMPLED, "resampled"))
        if exact_rates:
            log("  exact quadrature:")
            out["exact_quadrature_sweep"] = serving_sweep(srv, ids, exact, exact_rates, n_requests,
                                                          (sep.PREC_MIXED, "mixed"))
        out["largest_batch"] = int(srv.stats().largest_batch)
    rate = rates[0]
    reqs = resampled[:n_requests]
    arr = serving.poisson_arrivals(len(reqs), rate, seed=6000)
    with sep.Server(max_inflight=1, max_batch=1) as srv:
        ids = [srv.add_wedge(ph, sh) for ph, sh in holders]
        plain = serving.run(srv, ids, reqs, arr, force_precision=sep.PREC_FP64)
    out["one_at_a_time_fp64"] = {"poisson_rate_per_s": rate, "requests": n_requests,
                                 "throughput_req_per_s": round(plain["throughput"], 2),
                                 "p50_ms": round(plain["p50_ms"], 2), "p99_ms": round(plain["p99_ms"], 2)}
    return out


def timed_batch(pairs, step_sets, warmup_sets, precision, choice):
    """Eager run of the first set (kernel events around the persistent
    launch), then the warm-up sets and the timed sets through the graph-
    replayed batch, one fresh set per step; returns evals/s, per-batch
    latencies, elapsed seconds and the eager timing."""
    n_vec = sum(n for _, n in pairs)
    with sep.Batch(pairs, precision=precision, graph=False) as eager:
        eager.set_variant(choice["variant"], choice["blocks_per_cu"])
        eager.run(step_sets[0])
        eager.run(step_sets[0])
        tm = eager.timing()
    with sep.Batch(pairs, precision=precision, graph=True) as b:
        b.set_variant(choice["variant"], choice["blocks_per_cu"])
        for v in warmup_sets:
            b.run(v)
        sep.synchronize()
        lat = []
        t0 = time.perf_counter()
        for i, v in enumerate(step_sets):
            s0 = time.perf_counter()
            b.run(v)
            lat.append(time.perf_counter() - s0)
            log(f"  step {i}: batch of {n_vec} in {lat[-1] * 1e3:.1f} ms")
        sep.synchronize()
        elapsed = time.perf_counter() - t0
    return len(step_sets) * n_vec / elapsed, lat, elapsed, tm


def per_star_results(groups, vset, precision):
    out = []
    for (ctx, _, _), v in zip(groups, vset):
        out.extend(ctx.eval(np.asarray(v), precision=precision, per_star=True))
    return out


def accuracy(got, ref) -> dict:
    """Largest per-star |d ln P|, integral relative error and |d ln L| of
    `got` against `ref`, over every vector."""
    worst = {"per_star_dlnP": 0.0, "integral_rel": 0.0, "dlnL": 0.0}
    for a, b in zip(got, ref):
        worst["per_star_dlnP"] = max(worst["per_star_dlnP"],
                                     LN10 * float(np.max(np.abs(a.log10_p - b.log10_p))))
        ia = np.concatenate([[a.background_integral], a.stream_integrals])
        ib = np.concatenate([[b.background_integral], b.stream_integrals])
        worst["integral_rel"] = max(worst["integral_rel"], float(np.max(np.abs(ia - ib) / np.abs(ib))))
        worst["dlnL"] = max(worst["dlnL"], LN10 * abs(a.likelihood - b.likelihood))
    return {k: float(f"{v:.3e}") for k, v in worst.items()}


def tune(groups, vset, precision, args, label):
    pinned = args.variant if precision == sep.PREC_MIXED else args.resampled_variant
    if pinned is not None:
        return {"variant": pinned[0], "blocks_per_cu": pinned[1], "table": [], "s": 0.0,
                "scratch_bytes": None}
    t0 = time.perf_counter()
    choice = autotune.tune(groups, vset, precision, proxy_vectors=1, reps=1,
                           blocks_per_cu=(2, 4, 5), use_cache=False, log=log)
    choice["s"] = round(time.perf_counter() - t0, 2)
    log(f"autotuner ({label}): variant {choice['variant']} blocks/CU {choice['blocks_per_cu']}, "
        f"{choice['scratch_bytes']} B of spills per work-item ({choice['s']} s)")
    return choice


BATCH_DESC = ("batch: 64 parameter vectors (4 streams, 16 per wedge) over 4 wedges "
              "70x160x64 n=30 10k stars, 350x400x160 n=60 50k, 700x800x320 n=120 100k, "
              "700x1600x640 n=120 200k (r x mu x nu); one persistent launch replayed from a HIP "
              "graph; a fresh set of 64 vectors every step")


def bench_batch(args):
    t0 = time.perf_counter()
    groups, plist = build_batch()
    log(f"batch workload resident in {time.perf_counter() - t0:.2f} s")
    pairs = [(c, n) for c, n, _ in groups]
    n_vec = sum(n for _, n in pairs)
    steps, warmup = args.steps, args.warmup
    step_sets = [vector_set(m, plist) for m in range(steps)]
    warm_sets = [vector_set(steps + m, plist) for m in range(warmup)]

    # ---- headline: exact quadrature, mixed precision --------------------
    choice = tune(groups, step_sets[0], sep.PREC_MIXED, args, "exact quadrature, mixed")
    log("batched mixed precision on the reference's quadrature (headline)")
    value, lat, elapsed, tm = timed_batch(pairs, step_sets, warm_sets, sep.PREC_MIXED, choice)
    int_ms, points = tm.integral_ms, float(tm.integral_points)
    log(f"  eager: persistent integral launch {int_ms:.1f} ms of {tm.total_ms:.1f} ms, "
        f"{points:.4e} integrand evaluations")

    # ---- plain FP64, one at a time, the first set (the accuracy reference)
    log("plain FP64, one at a time")
    ref = []
    t0 = time.perf_counter()
    for (ctx, _, _), v in zip(groups, step_sets[0]):
        for row in v:
            ref.extend(ctx.eval(row, precision=sep.PREC_FP64, per_star=True))
    sep.synchronize()
    plain = n_vec / (time.perf_counter() - t0)
    log(f"  {plain:.3f} evals/s")
    log("accuracy of the headline against FP64, all 64 vectors")
    acc = accuracy(per_star_results(groups, step_sets[0], sep.PREC_MIXED), ref)
    log(f"  {acc}")

    roof = frozen_roofline(points, BATCH_STREAMS, int_ms)
    hw = measured_counters("integral_persistent_mixed_batch64")
    out = {
        "metric": "likelihood evals/s, batched mixed precision vs plain FP64 @ per-star |d lnP| <= 1e-7",
        "value": round(value, 4), "unit": "evals/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * elapsed / steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (seeded wedge geometries, parameter vectors and stars)",
        "config": {
            "workload": BATCH_DESC, "batch": n_vec, "precision": "mixed (MW_PREC_MIXED)",
            "quadrature": "exact: the reference's r-step x convolution points of every cell",
            "kernel_variant": choice["variant"], "blocks_per_cu": choice["blocks_per_cu"],
            "parallelism": "single GPU",
        },
        "accuracy_vs_fp64": dict(acc, bar_per_star_dlnP=1e-7,
                                 what="worst over all 64 vectors of the first timed set, against the "
                                      "GPU FP64 path (itself within 1e-12 of the oracle)"),
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 2),
                       "what": "one batch of 64 evaluations"},
        "plain_fp64": {"value": round(plain, 4), "unit": "evals/s",
                       "ratio_batched_mixed_over_plain_fp64": round(value / plain, 3),
                       "what": "the first set's 64 vectors, one evaluation at a time, FP64"},
        "roofline": dict(
            {"bound": "valu", "kernel": "integral_persistent (mixed, exact quadrature)"}, **roof,
            traffic=measured_traffic("integral_persistent_mixed_batch64"),
            avg_launch_ms=round(int_ms, 3),
            hardware=hw,
            note=("frac > 1 is the frozen model's over-charge, not a kernel above peak: the model "
                  "prices each stream's exponent at 22 ops and its exp at 20 flop-eq, while the "
                  "kernel spends 2 FP64 operations on the exponent and one v_exp_f32; the counters "
                  "(hardware) are the measure that depends on no model")),
        "tuning": {"seconds": choice["s"], "table": choice["table"],
                   "scratch_bytes_of_choice": choice["scratch_bytes"],
                   "rule": "fastest variant without register spills unless a spilling one is "
                           f"more than {autotune.SPILL_MARGIN:.0%} faster"},
    }

    # ---- the resampled quadrature, fresh vectors every step ---------------
    if not args.no_secondary:
        rs_choice = tune(groups, step_sets[0], sep.PREC_RESAMPLED, args, "resampled")
        log("resampled quadrature, fresh vectors every step")
        rs_steps = [vector_set(m, plist) for m in range(max(steps, 3))]
        rs_warm = [vector_set(1000 + m, plist) for m in range(max(warmup, 1))]
        v2, lat2, el2, tm2 = timed_batch(pairs, rs_steps, rs_warm, sep.PREC_RESAMPLED, rs_choice)
        acc2 = accuracy(per_star_results(groups, step_sets[0], sep.PREC_RESAMPLED), ref)
        log(f"  {v2:.1f} evals/s; accuracy {acc2}")
        out["resampled"] = {
            "value": round(v2, 4), "unit": "evals/s", "ratio_over_plain_fp64": round(v2 / plain, 3),
            "steps": len(rs_steps), "ms_per_step": round(1e3 * el2 / len(rs_steps), 3),
            "quadrature": ("resampled (MW_PREC_RESAMPLED): each region's r-step x convolution "
                           "weights folded by 16-node Lagrange interpolation onto grids sized to "
                           "the model's features; not the reference's quadrature"),
            "what": ("a fresh set of 64 vectors every timed step: re-plans, graph re-captures and "
                     "first folds of new grid levels are inside the timed region"),
            "accuracy_vs_fp64": dict(acc2, bar_per_star_dlnP=1e-7,
                                     what="worst over all 64 vectors of the first timed set"),
            "latency_ms": {"p50": round(1e3 * float(np.percentile(lat2, 50)), 2),
                           "p99": round(1e3 * float(np.percentile(lat2, 99)), 2)},
            "kernel_variant": rs_choice["variant"], "blocks_per_cu": rs_choice["blocks_per_cu"],
            "roofline": dict(
                {"bound": "valu", "kernel": "integral_persistent_split (mixed, resampled)"},
                **resampled_roofline(float(tm2.integral_points), float(tm2.integral_points_recurrence),
                                     BATCH_STREAMS, tm2.integral_ms,
                                     float(tm2.integral_points_background),
                                     float(tm2.integral_stream_terms)),
                traffic=measured_traffic("integral_persistent_resampled_batch64"),
                avg_launch_ms=round(tm2.integral_ms, 3),
                hardware=measured_counters("integral_persistent_resampled_batch64")),
        }
    for ctx, *_ in groups:
        ctx.close()

    if not args.no_full:
        log("full-size FP64 evaluation (3 warm-ups, 20 repetitions)")
        _, _, full = full_size_fp64(20, 3)
        out["full_size_fp64"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in full.items()}
        log("reference-size FP64 evaluation and 32 mixed vectors at full size")
        out["reference_size_fp64"] = reference_size_fp64()
        out["full_size_mixed_32"] = full_size_mixed_32()
        log(f"  {out['reference_size_fp64']['value']} evals/s; {out['full_size_mixed_32']['value']} evals/s")
        log("one GPU's share of the full-size evaluation split over 8 GPUs")
        out["one_gpu_share_of_8"] = one_gpu_share(full["ms_per_eval_median"])
        log(f"  {out['one_gpu_share_of_8']}")
    if not args.no_serving:
        log("serving run")
        out["serving"] = serving_run(args.serving_requests, args.serving_rates,
                                     args.serving_exact_rates, choice["variant"], choice["blocks_per_cu"])
    if not args.no_cpu:
        log("CPU baseline (oracle, one core and all cores)")
        from oracle import oracle as O
        out["cpu_baseline"] = cpu_baseline(plist, O.omp_threads())
    else:
        out["cpu_baseline"] = None
    return out


def bench_full(args):
    prec = {"fp64": sep.PREC_FP64, "mixed": sep.PREC_MIXED, "resampled": sep.PREC_RESAMPLED}[args.precision]
    p, stars, full = full_size_fp64(args.steps, args.warmup, prec)
    flops = full["flop_eq"]
    if prec == sep.PREC_FP64:
        roof = {"bound": "valu", "kernel": "integral_kernel", "achieved": full["achieved_tflops"],
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(full["achieved_tflops"] / FP64_PEAK_TFLOPS, 4),
                "traffic": measured_traffic("integral_kernel_fp64_full"),
                "algorithmic_flop_eq_per_launch": flops,
                "avg_launch_ms": full["integral_kernel_ms_median"],
                "hardware": measured_counters("integral_kernel_fp64_full")}
    elif prec == sep.PREC_MIXED:
        roof = dict({"bound": "valu", "kernel": "integral_kernel"},
                    **frozen_roofline(float(full["integrand_evaluations"]), synth.FULL_SIZE_STREAMS,
                                      full["integral_kernel_ms_median"]),
                    traffic=None, avg_launch_ms=full["integral_kernel_ms_median"])
    else:
        roof = dict({"bound": "valu", "kernel": "integral_kernel"},
                    **resampled_roofline(float(full["integrand_evaluations"]),
                                         float(full["integrand_evaluations_by_recurrence"]),
                                         synth.FULL_SIZE_STREAMS, full["integral_kernel_ms_median"],
                                         float(full["background_evaluations"]),
                                         float(full["integrand_evaluations_by_recurrence"])
                                         * synth.FULL_SIZE_STREAMS),
                    traffic=None, avg_launch_ms=full["integral_kernel_ms_median"])
    out = {
        "metric": "likelihood evals/s (separation integral + star pass)",
        "value": round(full["value"], 4), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(full["ms_per_eval_median"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == sep.PREC_FP64 else "f64+f32",
        "data": "synthetic (seeded wedge geometry, parameters and stars)",
        "integrand_evaluations": full["integrand_evaluations"],
        "config": {"workload": "full: " + full["workload"], "precision": args.precision,
                   "batch": 1, "parallelism": "single GPU"},
        "star_kernel_ms": full["star_kernel_ms_median"],
        "star_pass": full["star_pass"],
        "resident_setup_s": full["resident_setup_s"],
        "roofline": roof,
    }
    if not args.no_cpu:
        s, spent = cpu_seconds_per_eval(p, stars, 4, 4000, 0)
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
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the resampled batch (for profiling the headline alone)")
    ap.add_argument("--variant", type=int, nargs=2, metavar=("V", "BLOCKS_PER_CU"), default=None,
                    help="pin the headline batch's kernel variant and blocks per CU instead of "
                         "tuning (counter runs of one kernel: the tuner's timings move under "
                         "a profiler)")
    ap.add_argument("--resampled-variant", type=int, nargs=2, metavar=("V", "BLOCKS_PER_CU"), default=None,
                    help="pin the resampled batch's kernel variant and blocks per CU")
    ap.add_argument("--serving-requests", type=int, default=40)
    ap.add_argument("--serving-rates", type=float, nargs="+", default=[10.0, 50.0, 200.0, 800.0],
                    help="Poisson arrival rates (requests/s) of the serving sweep")
    ap.add_argument("--serving-exact-rates", type=float, nargs="*", default=[5.0, 20.0, 80.0],
                    help="Poisson arrival rates of the serving sweep on the exact quadrature "
                         "(none: skip it)")
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
