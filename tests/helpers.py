"""Shared builders for the tests: reference fixtures, synthetic workloads
and processes started one per rank."""
import json
import os
import socket
import subprocess
import sys
from pathlib import Path

import numpy as np

from milkyway_client_amd import params_io, synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder

GOLDEN = Path(__file__).resolve().parent / "golden"
REPO = Path(__file__).resolve().parent.parent


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


def free_port() -> int:
    with socket.socket(socket.AF_INET, socket.SOCK_STREAM) as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def run_ranks(cmd_tail, world: int, timeout: float = 180.0):
    """Starts one process per rank (the interpreter with cmd_tail, from the
    repository root, RANK / WORLD_SIZE / MASTER_* set for 127.0.0.1) and
    returns each rank's last printed JSON line, in rank order.  A rank that
    fails or outlives the timeout fails the test; every rank still running
    then is killed with its process group."""
    port = free_port()
    procs = []
    for r in range(world):
        env = dict(os.environ, RANK=str(r), LOCAL_RANK=str(r), WORLD_SIZE=str(world),
                   LOCAL_WORLD_SIZE=str(world), MASTER_ADDR="127.0.0.1", MASTER_PORT=str(port))
        print(f"starting rank {r}/{world}: {cmd_tail[:3]}", flush=True)
        procs.append(subprocess.Popen([sys.executable] + list(cmd_tail), env=env, cwd=REPO,
                                      stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                                      start_new_session=True))
    outs = []
    try:
        for r, p in enumerate(procs):
            out, err = p.communicate(timeout=timeout)
            assert p.returncode == 0, f"rank {r} exited with {p.returncode}:\n{err[-3000:]}"
            outs.append(json.loads(out.strip().splitlines()[-1]))
    finally:
        for p in procs:
            if p.poll() is None:
                os.killpg(p.pid, 9)
                p.wait()
    return outs
