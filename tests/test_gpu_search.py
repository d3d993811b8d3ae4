"""Search glue and checkpoints on the GPU.

* Sharded evaluation: each rank splits the regions' calculation ranges and
  the stars (split_astronomy_parameters, split_star_points), runs
  integral_f, the ranks' integrals are composed, each rank runs
  likelihood_f on the composed integrals, and likelihood_compose gives the
  client's value (astronomy_worker.c:35-134).  The ranks run one after the
  other on the one GPU.  Bar: the oracle's evaluate() - 3.0 within 1e-10,
  integrals within 1e-12 relative.
* Checkpoint resume: an integral pass stopped part-way through region 0 is
  written with mw_write_checkpoint, read into a fresh state and finished;
  the result must match an uninterrupted evaluation within the same bars.
* Hessian: mw_ctx_hessian (every stencil point in batched launches) against
  the same stencil applied to the oracle's likelihood values, and the
  errors program (errors.c:86-139) against mw_ctx_hessian + mw_errors.
"""
import ctypes as C
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN, ctest_vectors, fixture, holders, rel
from milkyway_client_amd import params_io
from milkyway_client_amd import separation as sep
from milkyway_client_amd.refstructs import (ASTRONOMY_PARAMETERS, EVALUATION_STATE, STAR_POINTS,
                                            ParamsHolder, StarsHolder)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module", autouse=True)
def lib():
    L = sep.load_library()
    sep.require_device()
    sep.set_precision(sep.PREC_FP64)
    return L


def c_inputs(L, fid):
    """Library-owned structs read from the fixture's files."""
    ap, sp, es = ASTRONOMY_PARAMETERS(), STAR_POINTS(), EVALUATION_STATE()
    assert L.mw_read_astronomy_parameters(str(GOLDEN / f"astronomy_parameters-{fid}-small.txt").encode(),
                                          C.byref(ap)) == 0
    assert L.mw_read_star_points(str(GOLDEN / f"stars-{fid}-first2000.txt").encode(), C.byref(sp)) == 0
    return ap, sp, es


@pytest.mark.parametrize("fid,ranks", [("86", 2), ("86", 3), ("11", 2)])
def test_sharded_evaluation_composes(lib, fid, ranks):
    vec = np.array(ctest_vectors()[fid], dtype=np.float64)
    S = None
    parts, owned = [], []
    for r in range(ranks):
        ap, sp, es = c_inputs(lib, fid)
        S = ap.number_streams
        lib.mw_split_astronomy_parameters(C.byref(ap), r, ranks)
        assert lib.mw_split_star_points(C.byref(sp), r, ranks) == 0
        assert lib.mw_initialize_state(C.byref(ap), C.byref(sp), C.byref(es)) == 0
        res = np.zeros(1 + S)
        print(f"rank {r}/{ranks}: integral_f", flush=True)
        lib.mw_hip_integral_f(vec.ctypes.data_as(_dp), res.ctypes.data_as(_dp), C.byref(ap),
                              C.byref(es), C.byref(sp))
        parts.append(res)
        owned.append((ap, sp, es))
    composed = np.zeros(1 + S)
    allparts = np.ascontiguousarray(np.stack(parts))
    lib.mw_integral_compose(allparts.ctypes.data_as(_dp), ranks, composed.ctypes.data_as(_dp),
                            C.byref(owned[0][0]))
    lres = []
    for r, (ap, sp, es) in enumerate(owned):
        out = np.zeros(2)
        lib.mw_hip_likelihood_f(composed.ctypes.data_as(_dp), out.ctypes.data_as(_dp), C.byref(ap),
                                C.byref(es), C.byref(sp))
        lres.append(out)
    lres = np.ascontiguousarray(np.stack(lres))
    got = lib.mw_likelihood_compose(lres.ctypes.data_as(_dp), ranks, 2000)

    p, stars = fixture(fid, n_stars=2000)
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars))
    print(f"composed={got!r} oracle={ref.likelihood - 3.0!r}")
    assert rel(composed[0], ref.background_integral) <= 1e-12
    assert rel(composed[1:], ref.stream_integrals) <= 1e-12
    assert abs(got - (ref.likelihood - 3.0)) <= 1e-10
    for ap, sp, es in owned:
        lib.mw_free_state(C.byref(es))
        lib.mw_free_star_points(C.byref(sp))
        lib.mw_free_astronomy_parameters(C.byref(ap))


def test_checkpoint_resume_of_integral_pass(lib, tmp_path):
    p, stars = fixture("86", n_stars=2000)
    vec = p.parameter_vector()
    ph, sh, es = holders(p, stars)
    a0, a1 = es.area(0), es.area(1)
    total0 = a0.r_steps * a0.mu_steps * a0.nu_steps
    total1 = a1.r_steps * a1.mu_steps * a1.nu_steps
    mid = total0 // 3 + 17
    # Part of region 0 only: stop at mid, leave region 1 untouched.
    a0.max_calculation = mid
    a1.current_calculation = total1
    sep.calculate_integrals(ph, es, sh)
    assert a0.current_calculation == mid
    # The state the reference would checkpoint mid-way through region 0.
    a0.max_calculation = total0
    a1.current_calculation = 0
    a1.background_integral = 0.0
    a1.stream_integrals[0] = 0.0
    es.es.current_integral = 0
    ck = tmp_path / "astronomy_checkpoint"
    assert lib.mw_write_checkpoint(str(ck).encode(), es.ptr) == 0
    assert 0 < lib.mw_calculate_progress(es.ptr) < 1

    # A fresh process: new state, read the checkpoint, finish.
    ph2, sh2, es2 = holders(p, stars)
    assert lib.mw_read_checkpoint(str(ck).encode(), es2.ptr) == 0
    assert es2.area(0).current_calculation == mid
    sep.calculate_integrals(ph2, es2, sh2)
    sep.calculate_likelihood(ph2, es2, sh2)
    ref = O.evaluate_full(ParamsHolder(p), StarsHolder(stars), vec)
    got = es2.es.prob_sum / len(stars)
    print(f"resumed lnL={got!r} oracle={ref.likelihood!r}")
    assert rel(es2.es.background_integral, ref.background_integral) <= 1e-12
    assert rel(es2.stream_integrals(), ref.stream_integrals) <= 1e-12
    assert abs(got - ref.likelihood) <= 1e-10


def stencil_points(x, h):
    n = len(x)
    pts = [x.copy()]
    for i in range(n):
        for sgn in (1, -1):
            v = x.copy()
            v[i] += sgn * h[i]
            pts.append(v)
    for i in range(n):
        for j in range(i + 1, n):
            for si, sj in ((1, 1), (1, -1), (-1, 1), (-1, -1)):
                v = x.copy()
                v[i] += si * h[i]
                v[j] += sj * h[j]
                pts.append(v)
    return pts


def stencil_hessian(vals, h):
    n = len(h)
    H = np.zeros((n, n))
    f0 = vals[0]
    for i in range(n):
        H[i, i] = (vals[1 + 2 * i] - 2 * f0 + vals[2 + 2 * i]) / h[i] ** 2
    k = 1 + 2 * n
    for i in range(n):
        for j in range(i + 1, n):
            c = vals[k:k + 4]
            H[i, j] = H[j, i] = (c[0] - c[1] - c[2] + c[3]) / (4 * h[i] * h[j])
            k += 4
    return H


def test_ctx_hessian_matches_oracle_values(lib):
    p, stars = fixture("79", n_stars=400, convolve=20)
    x = np.array(p.parameter_vector())
    ph = ParamsHolder(p)
    step = np.zeros(len(x))
    lib.mw_get_step(ph.ptr, step.ctypes.data_as(_dp))
    sh = StarsHolder(stars)
    with sep.Context(ph, sh) as ctx:
        H, g, v = ctx.hessian(x, step)
    pts = stencil_points(x, step)
    print("oracle stencil points:", len(pts), flush=True)
    vals = np.array([O.evaluate(ParamsHolder(p), sh, pt) - 3.0 for pt in pts])
    Href = stencil_hessian(vals, step)
    scale = np.max(np.abs(Href))
    print(f"max|dH|/max|H| = {np.max(np.abs(H - Href)) / scale:.2e}")
    assert abs(v - vals[0]) <= 1e-10
    assert np.max(np.abs(H - Href)) <= 1e-8 * scale
    gref = np.array([(vals[1 + 2 * i] - vals[2 + 2 * i]) / (2 * step[i]) for i in range(len(x))])
    assert np.max(np.abs(g - gref)) <= 1e-8 * max(1.0, np.max(np.abs(gref)))


def test_errors_program(lib, tmp_path):
    prog = sep.CLIENT_PATH.parent / "errors_hip"
    assert prog.exists(), "errors program not built"
    a = GOLDEN / "astronomy_parameters-79-small.txt"
    s = GOLDEN / "stars-79-first2000.txt"
    out = tmp_path / "errors.txt"
    print("errors program on fixture 79", flush=True)
    r = subprocess.run([str(prog), str(out), "-parameters", str(a), "-stars", str(s)],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr + r.stdout
    lines = out.read_text().splitlines()
    n = 8
    assert lines[0] == f"hessian [{n} x {n}]:"
    H_file = np.array([[float(t) for t in ln.split()] for ln in lines[1:1 + n]])
    assert lines[1 + n] == "" and lines[2 + n] == "SIGMA:"
    sd_file = np.array([float(t) for t in lines[-1].split()])
    assert lines[-2].startswith("Errors: q, r0, epsilon")

    p = params_io.read_astronomy_parameters(a)
    stars = params_io.read_star_points(s)
    ph, sh = ParamsHolder(p), StarsHolder(stars)
    step = np.zeros(n)
    lib.mw_get_step(ph.ptr, step.ctypes.data_as(_dp))
    with sep.Context(ph, sh) as ctx:
        H, _, _ = ctx.hessian(p.parameter_vector(), step)
    assert np.max(np.abs(H_file - H)) <= 1e-12 * np.max(np.abs(H))
    sigma, sd = sep.errors(H, len(stars))
    assert np.max(np.abs(sd_file - sd)) <= 5e-7 + 1e-6 * np.max(sd)
