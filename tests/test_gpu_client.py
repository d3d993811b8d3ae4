"""The command-line client on the GPU: the reference's CTest cases with values.

The reference registers its separation tests as runs of the client on a
parameter file, a star file and a parameter vector (tests/CMakeLists.txt:21-66)
and passes them when the client exits 0.  Here every small case whose files
ship with the reference (11, 12, 79, 86, each plain and -aux) runs the built
client as a subprocess, and the printed <search_likelihood> must equal the
oracle's evaluate() - 3.0 on the same files within 1e-10, and be bitwise the
in-process drop-in result (same library, same structs).  The star files are
the first 2000 stars of the reference's star files.

The -aux files optimise three more background parameters than the CTest
vectors hold, so the reference client refuses those runs with its
parameter-count error (src/milkyway_client_app.c:243-259); the client must
refuse them the same way, and is then run with the file's own full vector.
"""
import os
import re
import subprocess

import pytest

from helpers import GOLDEN, ctest_vectors, fixture
from milkyway_client_amd import separation as sep
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder, StateHolder
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LNL_TOL = 1e-10
MIXED_LNL_TOL = 1e-7
CASES = [(fid, aux) for fid in ("11", "12", "79", "86") for aux in (False, True)]
_LIKE = re.compile(r"<search_likelihood> (\S+) </search_likelihood>")


@pytest.fixture(scope="module", autouse=True)
def _device():
    sep.load_library()
    sep.require_device()
    assert sep.CLIENT_PATH.exists(), "client not built; run __graft_entry__.build()"


def files(fid, aux):
    a = GOLDEN / f"astronomy_parameters-{fid}-small{'-aux' if aux else ''}.txt"
    return a, GOLDEN / f"stars-{fid}-first2000.txt"


def run_client(fid, aux, precision=None, vec=None, expect_rc=0):
    a, p = files(fid, aux)
    env = dict(os.environ)
    env.pop("MW_HIP_PRECISION", None)
    if precision:
        env["MW_HIP_PRECISION"] = precision
    if vec is None:
        vec = ctest_vectors()[fid]
    cmd = [str(sep.CLIENT_PATH), "-p", str(p), "-a", str(a), *map(repr, map(float, vec))]
    print("client:", fid, "aux" if aux else "", precision or "fp64", len(vec), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == expect_rc, r.stderr
    if expect_rc:
        return None, r.stderr
    m = _LIKE.search(r.stderr)
    assert m, r.stderr
    assert "<search_application>" in r.stderr
    return float(m.group(1)), r.stderr


@pytest.mark.parametrize("fid,aux", CASES, ids=[f"{f}{'-aux' if a else ''}" for f, a in CASES])
def test_client_matches_oracle(fid, aux):
    p, stars = fixture(fid, n_stars=2000, aux=aux)
    vec = None
    if aux:
        _, err = run_client(fid, aux, expect_rc=1)
        assert "does not match the number of parameters" in err
        assert f"({p.optimized_parameter_count()})" in err
        vec = p.parameter_vector()
    got, _ = run_client(fid, aux, vec=vec)
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars)).likelihood - 3.0
    print(f"client={got!r} oracle={ref!r}")
    assert abs(got - ref) <= LNL_TOL

    # The in-process drop-in call on the same structs gives the same bits.
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    es = StateHolder(ph, sh)
    sep.set_precision(sep.PREC_FP64)
    assert sep.evaluate(ph, es, sh, p.parameter_vector()) - 3.0 == got


def test_client_mixed_precision():
    got, err = run_client("86", False, precision="mixed")
    assert " mixed </search_application>" in err
    p, stars = fixture("86", n_stars=2000)
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars)).likelihood - 3.0
    print(f"client mixed={got!r} oracle={ref!r} diff={got - ref:.3e}")
    assert abs(got - ref) <= MIXED_LNL_TOL
