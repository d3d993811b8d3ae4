#!/usr/bin/env python3
"""Benchmark of the MI355X separation likelihood evaluation.

A step is one likelihood evaluation (volume integral over every region plus
the star pass) of one parameter vector on the full-size wedge: main region
700 r x 1600 mu x 640 nu plus a 700 x 400 x 160 cut, 120 convolution points,
3 streams, 200k stars, seeded synthetic data.  Inputs (tables, stars) are
resident in HBM before the timed region; each step uploads its parameter
vector and reads back its results.

Prints one JSON line (the last line of stdout).  With --gpus above 1 it exits:
this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from milkyway_client_amd import separation as sep
from milkyway_client_amd import synth, workmodel
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder

FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector (and matrix) peak
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector peak (packed)

WORKLOADS = {
    "full": (synth.FULL_SIZE, synth.FULL_SIZE_STREAMS, 1),
    "reference": (synth.REFERENCE_SIZE, 1, 0),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p, stars, rows_main: int, rows_cut: int, n_star_sample: int):
    """Times the oracle (a C restatement of the reference's FP64 CPU path) on
    one host core over a bounded sample of the same workload: the first
    rows_main mu-rows of the main region, rows_cut of each cut, and the
    first n_star_sample stars; scales each part linearly to the full
    evaluation (work per mu-row and per star is uniform)."""
    from oracle import oracle as O
    O.build()
    ph = ParamsHolder(p)
    sample = stars[:n_star_sample]
    sh = StarsHolder(sample)
    st = StateHolder(ph, sh)
    full_scale = 0.0
    t0 = time.perf_counter()
    for g in range(ph.ap.number_integrals):
        a = st.area(g)
        rows = rows_main if g == 0 else rows_cut
        a.max_calculation = rows * a.nu_steps * a.r_steps
    O.integrals(ph, sh, st)
    t_int = time.perf_counter() - t0
    # every region's sample is rows/mu_steps of it; scale region by region
    # using the sampled points (regions have equal cost per point)
    sampled_pts = sum((rows_main if g == 0 else rows_cut) * r.nu_steps * r.r_steps
                      for g, r in enumerate(p.regions))
    total_pts = sum(r.mu_steps * r.nu_steps * r.r_steps for r in p.regions)
    full_scale = total_pts / sampled_pts
    st.es.background_integral = 1.0
    for s in range(p.number_streams):
        st.es.stream_integrals[s] = 1.0
    t0 = time.perf_counter()
    O.likelihood(ph, sh, st)
    t_star = time.perf_counter() - t0
    t_eval = t_int * full_scale + t_star * (len(stars) / len(sample))
    return {
        "value": 1.0 / t_eval, "unit": "evals/s", "cores": 1, "kind": "port",
        "sample": (f"oracle FP64, 1 thread: {rows_main} of {p.regions[0].mu_steps} main mu-rows"
                   + (f" + {rows_cut} of {p.regions[1].mu_steps} cut mu-rows" if len(p.regions) > 1 else "")
                   + f" ({t_int:.1f} s) and {len(sample)} of {len(stars)} stars ({t_star:.1f} s),"
                   f" scaled linearly to one full evaluation = {t_eval:.0f} s"),
        "seconds_per_eval": t_eval,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="full")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--batch", type=int, default=1, help="parameter vectors per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-rows", type=int, default=4)
    ap.add_argument("--cpu-stars", type=int, default=4000)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build evaluates on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    sep.load_library()
    sep.require_device()
    shape, n_streams, k = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    p, stars = synth.workload(k, shape, n_streams)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx = sep.Context(ph, sh, timing=True)
    t_setup = time.perf_counter() - t0
    log(f"workload {args.workload}: generated in {t_gen:.1f} s, resident setup {t_setup:.2f} s, "
        f"{ctx.resident_bytes / 2**20:.1f} MiB on device")
    prec = sep.PREC_MIXED if args.precision == "mixed" else sep.PREC_FP64
    vecs = synth.parameter_vectors(k, p, 4 * args.batch)

    def step(i):
        j = (i % 4) * args.batch
        return ctx.eval(vecs[j:j + args.batch], precision=prec)

    for i in range(args.warmup):
        step(i)
    sep.synchronize()
    lat, t_int, t_star = [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        res = step(i)
        lat.append(time.perf_counter() - s0)
        tm = ctx.timing()
        t_int.append(tm.integral_ms)
        t_star.append(tm.star_ms)
        log(f"step {i}: {lat[-1] * 1e3:.2f} ms (integral kernel {tm.integral_ms:.2f} ms, "
            f"star kernel {tm.star_ms:.2f} ms) lnL-3={res[0].likelihood - 3.0:.12f}")
    sep.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    evals_per_s = args.steps * args.batch / elapsed

    int_ms = float(np.mean(t_int))
    flops = workmodel.integral_flops(p.regions, p.convolve, n_streams) * args.batch
    peak = FP64_PEAK_TFLOPS if prec == sep.PREC_FP64 else FP32_PEAK_TFLOPS
    achieved = flops / (int_ms * 1e-3) / 1e12
    roofline = {
        "bound": "valu", "kernel": "integral_kernel", "achieved": round(achieved, 3),
        "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": None,
        "algorithmic_flop_eq_per_launch": flops,
        "avg_launch_ms": round(int_ms, 4),
    }
    out = {
        "metric": "likelihood evals/s (separation integral + star pass)",
        "value": round(evals_per_s, 4), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == sep.PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded wedge geometry, parameters and stars)",
        "config": {
            "workload": (f"{args.workload}: {' + '.join(f'{g.r_steps}x{g.mu_steps}x{g.nu_steps}' for g in p.regions)}"
                         f" (r x mu x nu), n_convolve={p.convolve}, streams={n_streams},"
                         f" stars={len(stars)}"),
            "precision": args.precision, "batch": args.batch,
            "parallelism": "single GPU",
        },
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "star_kernel_ms": round(float(np.mean(t_star)), 4),
        "resident_setup_s": round(t_setup, 3),
        "roofline": roofline,
    }
    if not args.no_cpu:
        cb = cpu_baseline(p, stars, args.cpu_rows, 1, args.cpu_stars)
        out["cpu_baseline"] = {k2: cb[k2] for k2 in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"]["value"] = round(cb["value"], 8)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
