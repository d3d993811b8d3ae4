"""Shared builders for the tests: reference fixtures and synthetic workloads."""
import json
from pathlib import Path

import numpy as np

from milkyway_client_amd import params_io, synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder

GOLDEN = Path(__file__).resolve().parent / "golden"


def ctest_vectors():
    return json.loads((GOLDEN / "ctest_parameter_vectors.json").read_text())


def fixture(fid: str, n_stars: int = 2000, aux: bool = False, convolve: int = None):
    """Reference test fixture fid (parameter file, first n_stars stars, and
    its CTest parameter vector applied)."""
    name = f"astronomy_parameters-{fid}-small{'-aux' if aux else ''}.txt"
    p = params_io.read_astronomy_parameters(GOLDEN / name)
    if convolve:
        p.convolve = convolve
    vec = ctest_vectors()[fid]
    if p.optimized_parameter_count() == len(vec):
        p.set_parameter_vector(vec)
    stars = params_io.read_star_points(GOLDEN / f"stars-{fid}-first2000.txt", limit=n_stars)
    return p, stars


def holders(p, stars):
    ph = ParamsHolder(p)
    sh = StarsHolder(stars)
    return ph, sh, StateHolder(ph, sh)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) if a.size else 0.0
