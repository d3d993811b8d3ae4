"""The C readers and the command-line client, on the CPU.

The library's readers (mw_read_astronomy_parameters, mw_read_star_points,
mw_initialize_state, ...) must fill the reference's structs exactly as the
Python readers do for every parameter and star file the reference's tests
use (tests/CMakeLists.txt:21-66), including their -aux variants, and reject
malformed files.  The client's argument handling and errors follow
src/milkyway_client_app.c:86-259; its GPU run is covered by the gpu tests.
No test here launches a kernel.
"""
import ctypes as C
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN, ctest_vectors
from milkyway_client_amd import params_io
from milkyway_client_amd import separation as sep
from milkyway_client_amd.refstructs import (ASTRONOMY_PARAMETERS, EVALUATION_STATE, STAR_POINTS,
                                            ParamsHolder, StarsHolder, StateHolder)

PARAM_FILES = sorted(GOLDEN.glob("astronomy_parameters-*.txt"))
STAR_FILES = sorted(GOLDEN.glob("stars-*.txt"))


@pytest.fixture(scope="module")
def lib():
    return sep.load_library()


def c_read_params(lib, path):
    ap = ASTRONOMY_PARAMETERS()
    rc = lib.mw_read_astronomy_parameters(str(path).encode(), C.byref(ap))
    return rc, ap


def same_params(a: ASTRONOMY_PARAMETERS, b: ASTRONOMY_PARAMETERS):
    """Field-by-field equality of two filled structs (doubles bitwise)."""
    nbg = a.number_background_parameters
    assert nbg == b.number_background_parameters
    for f in ("parameters_version", "background_weight", "number_streams",
              "number_stream_parameters", "convolve", "sgr_coordinates", "aux_bg_profile", "wedge",
              "number_integrals"):
        assert getattr(a, f) == getattr(b, f), f
    for f in ("background_parameters", "background_step", "background_min", "background_max",
              "background_optimize"):
        assert [getattr(a, f)[i] for i in range(nbg)] == [getattr(b, f)[i] for i in range(nbg)], f
    ns, nsp = a.number_streams, a.number_stream_parameters
    for f in ("stream_weights", "stream_weight_step", "stream_weight_min", "stream_weight_max",
              "stream_weight_optimize"):
        assert [getattr(a, f)[i] for i in range(ns)] == [getattr(b, f)[i] for i in range(ns)], f
    for f in ("stream_parameters", "stream_step", "stream_min", "stream_max", "stream_optimize"):
        for i in range(ns):
            assert ([getattr(a, f)[i][j] for j in range(nsp)]
                    == [getattr(b, f)[i][j] for j in range(nsp)]), f
    for g in range(a.number_integrals):
        ga, gb = a.integral[g][0], b.integral[g][0]
        for f, _ in ga._fields_:
            assert getattr(ga, f) == getattr(gb, f), (g, f)


@pytest.mark.parametrize("path", PARAM_FILES, ids=lambda p: p.name)
def test_c_parameter_reader_matches_python(lib, path):
    rc, ap = c_read_params(lib, path)
    assert rc == 0
    try:
        ph = ParamsHolder(params_io.read_astronomy_parameters(path))
        same_params(ap, ph.ap)
        assert lib.mw_get_optimized_parameter_count(C.byref(ap)) == \
            ph.py.optimized_parameter_count()
    finally:
        lib.mw_free_astronomy_parameters(C.byref(ap))


@pytest.mark.parametrize("path", STAR_FILES, ids=lambda p: p.name)
def test_c_star_reader_matches_python(lib, path):
    sp = STAR_POINTS()
    assert lib.mw_read_star_points(str(path).encode(), C.byref(sp)) == 0
    try:
        ref = params_io.read_star_points(path)
        assert sp.number_stars == len(ref)
        got = np.array([[sp.stars[i][k] for k in range(3)] for i in range(sp.number_stars)])
        assert np.array_equal(got, ref)
    finally:
        lib.mw_free_star_points(C.byref(sp))


def test_c_state_matches_python_state(lib):
    path = GOLDEN / "astronomy_parameters-86-small.txt"
    rc, ap = c_read_params(lib, path)
    assert rc == 0
    stars = params_io.read_star_points(GOLDEN / "stars-86-first2000.txt", limit=100)
    sh = StarsHolder(stars)
    es = EVALUATION_STATE()
    try:
        assert lib.mw_initialize_state(C.byref(ap), sh.ptr, C.byref(es)) == 0
        ref = StateHolder(ParamsHolder(params_io.read_astronomy_parameters(path)), sh)
        for f in ("current_integral", "number_streams", "number_integrals", "background_integral",
                  "current_star_point", "total_stars", "num_zero", "bad_jacobians", "prob_sum"):
            assert getattr(es, f) == getattr(ref.es, f), f
        for g in range(es.number_integrals):
            a, b = es.integral[g][0], ref.area(g)
            for f, _ in a._fields_:
                if f != "stream_integrals":
                    assert getattr(a, f) == getattr(b, f), (g, f)
            assert [a.stream_integrals[i] for i in range(es.number_streams)] == [0.0] * es.number_streams
    finally:
        lib.mw_free_state(C.byref(es))
        lib.mw_free_astronomy_parameters(C.byref(ap))


def test_c_set_parameters_order(lib):
    """set_astronomy_parameters order (parameters.c:332-363): background
    first, then each stream's weight and parameters, skipping fixed ones."""
    path = GOLDEN / "astronomy_parameters-11-small.txt"
    rc, ap = c_read_params(lib, path)
    assert rc == 0
    try:
        vec = np.array(ctest_vectors()["11"], dtype=np.float64)
        n = lib.mw_get_optimized_parameter_count(C.byref(ap))
        assert n == len(vec)
        lib.mw_set_astronomy_parameters(C.byref(ap), vec.ctypes.data_as(C.POINTER(C.c_double)))
        p = params_io.read_astronomy_parameters(path)
        p.set_parameter_vector(vec)
        same_params(ap, ParamsHolder(p).ap)
    finally:
        lib.mw_free_astronomy_parameters(C.byref(ap))


def test_writer_round_trip_and_version_001(lib, tmp_path):
    """A file written in the reference's writer format (parameters.c:230-305)
    reads back identically; a file without a version line is version 0.01
    and has no aux_bg_profile line (parameters.c:111-117, :157-160)."""
    p = params_io.read_astronomy_parameters(GOLDEN / "astronomy_parameters-79-small-aux.txt")
    out = tmp_path / "rt.txt"
    params_io.write_astronomy_parameters(out, p)
    rc, ap = c_read_params(lib, out)
    assert rc == 0
    try:
        same_params(ap, ParamsHolder(params_io.read_astronomy_parameters(out)).ap)
        assert ap.aux_bg_profile == 1
    finally:
        lib.mw_free_astronomy_parameters(C.byref(ap))

    text = out.read_text().splitlines()
    old = [ln for ln in text if not ln.startswith(("parameters_version", "aux_bg_profile"))]
    out.write_text("\n".join(old) + "\n")
    rc, ap = c_read_params(lib, out)
    assert rc == 0
    try:
        assert ap.parameters_version == 0.01 and ap.aux_bg_profile == 0
        assert ap.number_integrals == len(p.regions)
    finally:
        lib.mw_free_astronomy_parameters(C.byref(ap))


@pytest.mark.parametrize("path", PARAM_FILES, ids=lambda p: p.name)
def test_c_writer_matches_python_writer(lib, path, tmp_path):
    """Both writers emit the reference's format (parameters.c:230-305) byte
    for byte, and what the C writer wrote reads back to the same struct."""
    rc, ap = c_read_params(lib, path)
    assert rc == 0
    try:
        out = tmp_path / "c.txt"
        assert lib.mw_write_astronomy_parameters(str(out).encode(), C.byref(ap)) == 0
        py = tmp_path / "py.txt"
        params_io.write_astronomy_parameters(py, params_io.read_astronomy_parameters(path))
        assert out.read_text() == py.read_text()
        rc2, back = c_read_params(lib, out)
        assert rc2 == 0
        same_params(back, ParamsHolder(params_io.read_astronomy_parameters(out)).ap)
        lib.mw_free_astronomy_parameters(C.byref(back))
    finally:
        lib.mw_free_astronomy_parameters(C.byref(ap))


def test_c_star_writer_matches_python_writer(lib, tmp_path):
    stars = params_io.read_star_points(GOLDEN / "stars-79-first2000.txt", limit=50)
    sh = StarsHolder(stars)
    out = tmp_path / "c.txt"
    assert lib.mw_write_star_points(str(out).encode(), sh.ptr) == 0
    py = tmp_path / "py.txt"
    params_io.write_star_points(py, stars)
    assert out.read_text() == py.read_text()


def test_commas_are_optional_in_arrays(lib, tmp_path):
    """io_util.c:63-67: the separator directive matches zero or one comma."""
    src = (GOLDEN / "astronomy_parameters-86-small.txt").read_text()
    out = tmp_path / "nocomma.txt"
    out.write_text(src.replace("stream_step[5]: 0.30000000, 0.30000000,",
                               "stream_step[5]: 0.30000000  0.30000000"))
    assert out.read_text() != src
    rc, ap = c_read_params(lib, out)
    assert rc == 0
    lib.mw_free_astronomy_parameters(C.byref(ap))


@pytest.mark.parametrize("mutate", [
    lambda s: s.replace("background_parameters[4]: 1.0000000000,", "background_parameters[4]: x,"),
    lambda s: s.replace("background_parameters[4]", "background_parameters[3]"),
    lambda s: s.replace("number_streams: 1, 5", "number_streams: 1, 4"),
    lambda s: s.replace("stream_parameters[5]", "stream_parameters[3]"),
    lambda s: s.replace("convolve: 120", "convolve: -1"),
    lambda s: s.replace("mu[min,max,steps]: 310.000000, 420.000000, 160", "mu[min,max,steps]: 310.0"),
    lambda s: s.replace("number_cuts: 1", "number_cuts: 2"),
    lambda s: s.split("wedge:")[0],
], ids=["bad-number", "short-bg-array", "stream-params-not-5", "short-stream-array",
        "negative-convolve", "truncated-range", "missing-cut", "truncated-file"])
def test_malformed_parameter_files_are_rejected(lib, tmp_path, mutate):
    src = (GOLDEN / "astronomy_parameters-86-small.txt").read_text()
    bad = mutate(src)
    assert bad != src
    out = tmp_path / "bad.txt"
    out.write_text(bad)
    rc, ap = c_read_params(lib, out)
    assert rc == 1
    assert ap.background_parameters is None or not bool(ap.background_parameters)


def test_bad_star_files_are_rejected(lib, tmp_path):
    sp = STAR_POINTS()
    assert lib.mw_read_star_points(str(tmp_path / "absent.txt").encode(), C.byref(sp)) == 1
    short = tmp_path / "short.txt"
    short.write_text("3\n1 2 3\n4 5 6\n")
    assert lib.mw_read_star_points(str(short).encode(), C.byref(sp)) == 1
    assert sp.number_stars == 0
    empty = tmp_path / "empty.txt"
    empty.write_text("0\n")
    assert lib.mw_read_star_points(str(empty).encode(), C.byref(sp)) == 0
    assert sp.number_stars == 0
    lib.mw_free_star_points(C.byref(sp))


def run_client(*args, env=None):
    assert sep.CLIENT_PATH.exists(), "client not built; run __graft_entry__.build()"
    return subprocess.run([str(sep.CLIENT_PATH), *map(str, args)], capture_output=True, text=True,
                          timeout=60, env=env)


def test_client_rejects_wrong_parameter_count():
    r = run_client("-p", GOLDEN / "stars-86-first2000.txt", "-a",
                   GOLDEN / "astronomy_parameters-86-small.txt", 1, 2, 3)
    assert r.returncode == 1
    assert "number of parameters from the command line (3)" in r.stderr
    assert "(8)" in r.stderr


def test_client_usage_errors():
    r = run_client("-p", GOLDEN / "stars-86-first2000.txt", "-a",
                   GOLDEN / "astronomy_parameters-86-small.txt")
    assert r.returncode == 1 and "Could not parse parameters" in r.stderr
    r = run_client("-x", "1")
    assert r.returncode == 1 and "bad option" in r.stderr
    r = run_client("--help")
    assert r.returncode == 0 and "--astronomy-parameter-file" in r.stdout
    r = run_client("-a", "absent-a.txt", "-p", "absent-p.txt", "1")
    assert r.returncode == 1
    assert "APP: error reading astronomy parameters from file absent-a.txt" in r.stderr
    assert "APP: error reading star points from file absent-p.txt" in r.stderr


def test_client_without_gpu_fails_loudly():
    """With no GPU the evaluator exits with its error code (no CPU fallback)."""
    if sep.device_count() > 0:
        pytest.skip("a GPU is visible; the gpu tests run the client")
    vec = ctest_vectors()["86"]
    r = run_client("-p", GOLDEN / "stars-86-first2000.txt", "-a",
                   GOLDEN / "astronomy_parameters-86-small.txt", *vec)
    assert r.returncode == 4
    assert "APP: error calculating" in r.stderr
    assert "<search_likelihood>" not in r.stderr
