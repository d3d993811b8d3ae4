"""Writes the full-size oracle fixture: one whole evaluation of the
700 x 1600 x 640 wedge plus its 700 x 400 x 160 cut (n = 120, 3 streams,
200k stars) by the CPU oracle.

The workload is synth.workload(1, FULL_SIZE, 3) with its own parameter
values.  The integrals come from the oracle's OpenMP mode (mu rows over
threads, per-row Kahan sums combined in row order), the star pass from the
serial oracle, which also yields every star's log10 P.  Written next to this
script:

  full_size_oracle.json        region integrals, the combined integrals,
                               prob_sum, num_zero, the likelihood, the
                               background-only and stream-only means, and
                               the sum of every star's log10 P
  full_size_log10p_every10.npy log10 P of stars 0, 10, 20, ... (20,000 values)

Run once in the build container (about 5 minutes on 8 cores):
    python tests/golden/make_full_size_fixture.py [threads]
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
sys.path.insert(0, str(REPO))

from milkyway_client_amd import synth  # noqa: E402
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder  # noqa: E402
from oracle import oracle as O  # noqa: E402

STRIDE = 10


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else O.omp_threads()
    O.build()
    p, stars = synth.workload(1, synth.FULL_SIZE, synth.FULL_SIZE_STREAMS)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    st = StateHolder(ph, sh)
    t0 = time.perf_counter()
    O.integrals(ph, sh, st, threads=threads)
    t_int = time.perf_counter() - t0
    print(f"integrals: {t_int:.1f} s on {threads} threads", flush=True)
    t0 = time.perf_counter()
    res = O.likelihood(ph, sh, st)
    t_star = time.perf_counter() - t0
    print(f"star pass: {t_star:.1f} s (serial)", flush=True)
    lp = np.asarray(res.log10_p, dtype=np.float64)
    out = {
        "workload": "synth.workload(1, synth.FULL_SIZE, 3) with its own parameter values",
        "parameter_vector": [float(x) for x in p.parameter_vector()],
        "regions": [[float(bg)] + [float(x) for x in sti] for bg, sti in res.regions],
        "background_integral": float(res.background_integral),
        "stream_integrals": [float(x) for x in res.stream_integrals],
        "prob_sum": float(res.prob_sum),
        "num_zero": int(res.num_zero),
        "likelihood": float(res.likelihood),
        "bg_only_mean": float(res.bg_only_mean),
        "st_only_mean": [float(x) for x in res.st_only_mean],
        "sum_log10p_all_stars": float(np.sum(lp)),
        "n_stars": int(len(lp)),
        "log10p_stride": STRIDE,
        "oracle_seconds": {"integrals": round(t_int, 1), "integral_threads": threads,
                           "star_pass_serial": round(t_star, 1)},
    }
    (HERE / "full_size_oracle.json").write_text(json.dumps(out, indent=1) + "\n")
    np.save(HERE / "full_size_log10p_every10.npy", lp[::STRIDE].copy())
    print(json.dumps({k: out[k] for k in ("background_integral", "stream_integrals", "likelihood")}))


if __name__ == "__main__":
    main()
