"""Host-side search glue of the library (mw_search.h), on the CPU.

The central-difference Hessian is checked on polynomials whose Hessian is
known in closed form (the stencil is exact for them up to rounding), the LUP
inverse against numpy.linalg.inv, the compose steps against their defining
sums, and the splitters by the partition they must form.  No test here
launches a kernel.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import GOLDEN
from milkyway_client_amd import params_io
from milkyway_client_amd import separation as sep
from milkyway_client_amd.refstructs import ParamsHolder

_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def lib():
    return sep.load_library()


def arr(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def hessian_of(lib, n, point, step, f):
    calls = []

    @sep.VECTOR_EVAL_FN
    def cb(user, vectors, count, values):
        v = np.ctypeslib.as_array(vectors, shape=(count, n))
        out = np.ctypeslib.as_array(values, shape=(count,))
        out[:] = [f(row) for row in v]
        calls.append(count)
        return 0

    H = np.zeros((n, n))
    g = np.zeros(n)
    val = C.c_double()
    x, h = arr(point), arr(step)
    rc = lib.mw_hessian(n, x.ctypes.data_as(_dp), h.ctypes.data_as(_dp), cb, None,
                        H.ctypes.data_as(_dp), g.ctypes.data_as(_dp), C.byref(val))
    return rc, H, g, val.value, calls


@pytest.mark.parametrize("n", [1, 2, 5, 8])
def test_hessian_exact_on_polynomials(lib, n):
    """f = 1/2 x'Ax + b'x + c + sum d_i x_i^3: H = A + diag(6 d x); the
    stencil's error terms involve fourth derivatives, which vanish."""
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    A = A + A.T
    b, d = rng.normal(size=n), rng.normal(size=n)
    x = rng.normal(size=n)
    h = rng.uniform(0.05, 0.2, size=n)

    def f(v):
        return 0.5 * v @ A @ v + b @ v + 1.25 + np.sum(d * v ** 3)

    rc, H, g, val, calls = hessian_of(lib, n, x, h, f)
    assert rc == 0
    assert calls == [2 * n * n + 1]          # one batched call with every stencil point
    want = A + np.diag(6 * d * x)
    assert np.max(np.abs(H - want)) < 1e-9 * max(1.0, np.max(np.abs(want)))
    assert np.array_equal(H, H.T)
    # gradient: exact for the quadratic part, h^2 d_i off for the cubic one
    g_want = A @ x + b + 3 * d * x ** 2 + d * h ** 2
    assert np.max(np.abs(g - g_want)) < 1e-9
    assert abs(val - f(x)) < 1e-12


def test_hessian_rejects_zero_steps_and_propagates_errors(lib):
    rc, *_ = hessian_of(lib, 2, [0.0, 0.0], [0.1, 0.0], lambda v: 0.0)
    assert rc == 1

    @sep.VECTOR_EVAL_FN
    def bad(user, vectors, count, values):
        return 7

    H = np.zeros((2, 2))
    x = arr([0.0, 0.0])
    h = arr([0.1, 0.1])
    assert lib.mw_hessian(2, x.ctypes.data_as(_dp), h.ctypes.data_as(_dp), bad, None,
                          H.ctypes.data_as(_dp), None, None) == 7


def test_matrix_invert_against_numpy(lib):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 8, 20):
        a = arr(rng.normal(size=(n, n)) + n * np.eye(n))
        if n > 1:
            a[0, 0] = 0.0                       # forces a row exchange
        inv = np.zeros((n, n))
        assert lib.mw_matrix_invert(n, a.ctypes.data_as(_dp), inv.ctypes.data_as(_dp)) == 0
        assert np.max(np.abs(inv - np.linalg.inv(a))) < 1e-12 * np.max(np.abs(np.linalg.inv(a)))
    singular = arr([[1.0, 2.0], [2.0, 4.0]])
    inv = np.zeros((2, 2))
    assert lib.mw_matrix_invert(2, singular.ctypes.data_as(_dp), inv.ctypes.data_as(_dp)) == 1


def test_errors_are_scaled_inverse(lib):
    """errors.c:38-84: sigma = H^-1 / N, stddev = sqrt(|diag sigma|)."""
    rng = np.random.default_rng(4)
    m = rng.normal(size=(6, 6))
    H = -(m @ m.T + 6 * np.eye(6))              # a maximum: negative definite
    sigma, sd = sep.errors(H, 2000)
    assert np.max(np.abs(sigma - np.linalg.inv(H) / 2000)) < 1e-15
    assert np.allclose(sd, np.sqrt(np.abs(np.diag(np.linalg.inv(H) / 2000))), rtol=1e-13)
    with pytest.raises(np.linalg.LinAlgError):
        sep.errors(np.zeros((3, 3)), 10)


def test_compose(lib):
    """integral_compose sums the ranks' (1 + S)-vectors; likelihood_compose is
    sum(prob_sum) / (N - sum(bad_jacobians)) - 3 (astronomy_worker.c:68-134)."""
    p = params_io.read_astronomy_parameters(GOLDEN / "astronomy_parameters-11-small.txt")
    ph = ParamsHolder(p)
    S = p.number_streams
    parts = np.random.default_rng(1).uniform(size=(5, 1 + S))
    out = np.zeros(1 + S)
    lib.mw_integral_compose(arr(parts).ctypes.data_as(_dp), 5, out.ctypes.data_as(_dp), ph.ptr)
    assert np.allclose(out, parts.sum(axis=0), rtol=1e-15)
    res = arr([[-1000.5, 0.0], [-2000.25, 2.0], [-500.0, 1.0]])
    got = lib.mw_likelihood_compose(res.ctypes.data_as(_dp), 3, 1203)
    assert got == pytest.approx((-1000.5 - 2000.25 - 500.0) / (1203 - 3.0) - 3.0, rel=1e-15)


@pytest.mark.parametrize("ranks", [1, 2, 3, 7])
def test_split_ranges_partition_every_region(lib, ranks):
    p = params_io.read_astronomy_parameters(GOLDEN / "astronomy_parameters-86-small.txt")
    spans = []
    for r in range(ranks):
        ph = ParamsHolder(p)
        lib.mw_split_astronomy_parameters(ph.ptr, r, ranks)
        spans.append([(ph.region(g).min_calculation, ph.region(g).max_calculation)
                      for g in range(len(p.regions))])
    for g, reg in enumerate(p.regions):
        total = reg.r_steps * reg.mu_steps * reg.nu_steps
        cur = 0
        for r in range(ranks):
            lo, hi = spans[r][g]
            assert lo == cur and hi >= lo
            cur = hi
        assert cur == total


@pytest.mark.parametrize("ranks", [1, 2, 3, 7])
def test_split_stars_partition(lib, ranks):
    from milkyway_client_amd.refstructs import STAR_POINTS
    pieces = []
    for r in range(ranks):
        # a library-owned copy, so that the split can free its rows
        sp = STAR_POINTS()
        path = GOLDEN / "stars-79-first2000.txt"
        assert lib.mw_read_star_points(str(path).encode(), C.byref(sp)) == 0
        keep = sp.number_stars
        assert lib.mw_split_star_points(C.byref(sp), r, ranks) == 0
        pieces.append(np.array([[sp.stars[i][k] for k in range(3)] for i in range(sp.number_stars)]))
        lib.mw_free_star_points(C.byref(sp))
        assert keep == 2000
    full = params_io.read_star_points(GOLDEN / "stars-79-first2000.txt")
    assert np.array_equal(np.concatenate([x for x in pieces if len(x)]), full)


def test_vector_accessors(lib):
    """get_search_parameters / get_step / get_min / get_max in the vector order
    of set_astronomy_parameters (parameters.c:332-500)."""
    p = params_io.read_astronomy_parameters(GOLDEN / "astronomy_parameters-11-small.txt")
    ph = ParamsHolder(p)
    n = p.optimized_parameter_count()
    want = {"mw_get_search_parameters": [], "mw_get_step": [], "mw_get_min_parameters": [],
            "mw_get_max_parameters": []}
    for i, opt in enumerate(p.background_optimize):
        if opt:
            want["mw_get_search_parameters"].append(p.background_parameters[i])
            want["mw_get_step"].append(p.background_step[i])
            want["mw_get_min_parameters"].append(p.background_min[i])
            want["mw_get_max_parameters"].append(p.background_max[i])
    for s in p.streams:
        if s.weight_optimize:
            want["mw_get_search_parameters"].append(s.weight)
            want["mw_get_step"].append(s.weight_step)
            want["mw_get_min_parameters"].append(s.weight_min)
            want["mw_get_max_parameters"].append(s.weight_max)
        for j, opt in enumerate(s.optimize):
            if opt:
                want["mw_get_search_parameters"].append(s.parameters[j])
                want["mw_get_step"].append(s.step[j])
                want["mw_get_min_parameters"].append(s.min[j])
                want["mw_get_max_parameters"].append(s.max[j])
    for name, w in want.items():
        out = np.zeros(n)
        getattr(lib, name)(ph.ptr, out.ctypes.data_as(_dp))
        assert list(out) == w, name
    assert np.array_equal(want["mw_get_search_parameters"], p.parameter_vector())
