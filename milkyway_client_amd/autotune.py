"""Autotuner of the batched persistent launch.

For a batch signature (each wedge's r x mu x nu grid, convolution points,
stream count, star count and vector count, plus the precision) it times
every kernel variant (inner-loop unroll 2/4 x occupancy target default/4/5
waves per EU) at several persistent grid sizes (1, 2, 4, 5 or 8 workgroups per
compute unit) and caches its choice in a JSON file: the fastest of the
variants whose kernels spill no registers to scratch, unless a spilling one
is more than SPILL_MARGIN faster (spills cost scratch traffic that the
proxy's timing hardly shows).

Timing runs on a proxy of the batch: the same resident wedges with at most
``proxy_vectors`` vectors each, so tuning costs a small fraction of one full
batch.  Every variant computes bitwise-identical results (the tests check
this), so the choice only affects speed.
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Dict, List, Sequence, Tuple

from . import separation as sep

DEFAULT_CACHE = Path(os.environ.get("MW_TUNING_CACHE",
                                    Path.home() / ".cache" / "mw_separation" / "tuning.json"))
BLOCKS_PER_CU = (1, 2, 4, 5, 8)
# Bumped whenever the library's table of kernel variants changes meaning
# (variant v = unroll {2, 4} x occupancy target {compiler's, 4, 5} waves),
# so that cached choices of an older table are not reused.
VARIANT_TABLE = 2
SPILL_MARGIN = 0.01


def signature(groups, precision: int) -> str:
    """groups: list of (Context, n_vectors, shape-description tuple)."""
    parts = sorted(f"{desc}x{n}" for _, n, desc in groups)
    return f"t{VARIANT_TABLE}p{precision}:" + ";".join(parts)


def load_cache(path: Path = DEFAULT_CACHE) -> Dict:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def save_cache(cache: Dict, path: Path = DEFAULT_CACHE) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
    os.replace(tmp, path)


def tune(groups, vectors: Sequence, precision: int, proxy_vectors: int = 1, reps: int = 2,
         blocks_per_cu: Sequence[int] = BLOCKS_PER_CU, cache_path: Path = DEFAULT_CACHE,
         use_cache: bool = True, log=None) -> Dict:
    """groups: list of (Context, n_vectors, shape-description); vectors: one
    (n_vectors, n_params) array per group.  Returns the chosen
    {"variant", "blocks_per_cu", "ms", "table"}."""
    sig = signature(groups, precision)
    cache = load_cache(cache_path) if use_cache else {}
    if sig in cache:
        return cache[sig]
    proxy = [(ctx, min(n, proxy_vectors)) for ctx, n, _ in groups]
    pvecs = [v[:n] for (_, n), v in zip(proxy, vectors)]
    table: List[Tuple[int, int, float]] = []
    spills: Dict[int, int] = {}
    with sep.Batch(proxy, precision=precision, graph=True) as b:
        for variant in range(sep.Batch.variant_count()):
            for bpc in blocks_per_cu:
                b.set_variant(variant, bpc)
                b.run(pvecs)                       # capture + warm
                best = float("inf")
                for _ in range(reps):
                    b.run(pvecs)
                    best = min(best, b.timing().total_ms)
                table.append((variant, bpc, best))
                if log:
                    log(f"tune variant {variant} blocks/CU {bpc}: {best:.3f} ms")
            spills[variant] = b.variant_scratch(variant)
    v, bpc, ms = pick(table, spills)
    choice = {"variant": v, "blocks_per_cu": bpc, "ms": ms,
              "table": [list(t) for t in table], "scratch_bytes": spills[v],
              "tuned_at": time.time()}
    if use_cache:
        cache[sig] = choice
        save_cache(cache, cache_path)
    return choice


def pick(table: Sequence[Tuple[int, int, float]], spills: Dict[int, int],
         margin: float = SPILL_MARGIN) -> Tuple[int, int, float]:
    """The fastest (variant, blocks_per_cu, ms) of table among variants that
    spill nothing (spills[v] == 0), unless the fastest overall is more than
    ``margin`` faster than that."""
    best = min(table, key=lambda t: t[2])
    clean = [t for t in table if spills.get(t[0], 0) == 0]
    if clean:
        c = min(clean, key=lambda t: t[2])
        if c[2] <= best[2] * (1.0 + margin):
            return c
    return best
