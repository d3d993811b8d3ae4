"""ctypes wrapper of the CPU oracle (test infrastructure only).

Wraps ``liboracle.so`` built from ``mw_oracle.c`` by ``oracle/Makefile``.
Every entry point restates a reference function; see the citations in
``mw_oracle.c``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from milkyway_client_amd.refstructs import (ASTRONOMY_PARAMETERS, EVALUATION_STATE,
                                            STAR_POINTS, ParamsHolder, StarsHolder,
                                            StateHolder)

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_dp = C.POINTER(C.c_double)
_lib = None


def build(force: bool = False) -> Path:
    """Compile the oracle with oracle/Makefile."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)] + (["-B"] if force else []),
                       check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        AP, ES, SP = (C.POINTER(ASTRONOMY_PARAMETERS), C.POINTER(EVALUATION_STATE),
                      C.POINTER(STAR_POINTS))
        L.orc_gauss_legendre.argtypes = [C.c_double, C.c_double, _dp, _dp, C.c_int]
        L.orc_gc_to_lb.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_gc_to_eq.argtypes = [C.c_double, C.c_double, _dp, _dp, C.c_double, C.c_double]
        L.orc_eq_to_gal.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_gc_to_sgr.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_sgr_to_gal.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_wedge_incl.argtypes = [C.c_int]
        L.orc_wedge_incl.restype = C.c_double
        L.orc_get_node.restype = C.c_double
        L.orc_lbr2xyz.argtypes = [_dp, _dp]
        L.orc_r_constants.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_double, C.c_double,
                                      _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_point_probabilities.argtypes = [AP, C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.orc_point_probabilities.restype = C.c_int
        for f in ("orc_calculate_integrals", "orc_calculate_likelihood"):
            getattr(L, f).argtypes = [AP, ES, SP]
            getattr(L, f).restype = C.c_int
        L.orc_calculate_likelihood_ex.argtypes = [AP, ES, SP, _dp, _dp, _dp]
        L.orc_calculate_likelihood_ex.restype = C.c_int
        L.orc_evaluate.argtypes = [_dp, AP, ES, SP]
        L.orc_evaluate.restype = C.c_double
        L.orc_get_optimized_parameter_count.argtypes = [AP]
        L.orc_get_optimized_parameter_count.restype = C.c_int
        L.orc_set_astronomy_parameters.argtypes = [AP, _dp]
        L.orc_reset_evaluation_state.argtypes = [ES]
        L.orc_calculate_integrals_omp.argtypes = [AP, ES, C.c_int]
        L.orc_calculate_integrals_omp.restype = C.c_int
        L.orc_calculate_likelihood_omp.argtypes = [AP, ES, SP, C.c_int]
        L.orc_calculate_likelihood_omp.restype = C.c_int
        L.orc_stripe_normal.argtypes = [C.c_int, _dp]
        L.orc_sgr_stripe_normal.argtypes = [C.c_int, _dp]
        L.orc_get_transform.argtypes = [_dp, _dp, _dp]
        L.orc_transform_point.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_prob_ok.argtypes = [C.c_int, _dp, C.c_double]
        L.orc_prob_ok.restype = C.c_int
        L.orc_separation.argtypes = [AP, SP, C.c_double, _dp, C.c_ulong, C.POINTER(C.c_int), _dp,
                                     _dp, _dp, C.POINTER(C.c_int)]
        L.orc_separation.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def gauss_legendre(n: int, x1: float = -1.0, x2: float = 1.0):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(x1, x2, _ptr(x), _ptr(w), n)
    return x, w


def gc_to_lb(wedge: int, sgr: int, mu: float, nu: float):
    l, b = C.c_double(), C.c_double()
    lib().orc_gc_to_lb(wedge, sgr, mu, nu, C.byref(l), C.byref(b))
    return l.value, b.value


def gc_to_eq(mu, nu, node, inc):
    a, d = C.c_double(), C.c_double()
    lib().orc_gc_to_eq(mu, nu, C.byref(a), C.byref(d), node, inc)
    return a.value, d.value


def eq_to_gal(ra, dec):
    l, b = C.c_double(), C.c_double()
    lib().orc_eq_to_gal(ra, dec, C.byref(l), C.byref(b))
    return l.value, b.value


def gc_to_sgr(mu, nu, wedge):
    a, b = C.c_double(), C.c_double()
    lib().orc_gc_to_sgr(mu, nu, wedge, C.byref(a), C.byref(b))
    return a.value, b.value


def sgr_to_gal(lam, beta):
    l, b = C.c_double(), C.c_double()
    lib().orc_sgr_to_gal(lam, beta, C.byref(l), C.byref(b))
    return l.value, b.value


def lbr2xyz(l, b, r):
    lbr = np.array([l, b, r], dtype=np.float64)
    xyz = np.zeros(3)
    lib().orc_lbr2xyz(_ptr(lbr), _ptr(xyz))
    return xyz


def r_constants(n: int, r_steps: int, r_min: float, r_step: float, mu_step: float,
                nu_steps: int, nu_min: float, nu_step: float):
    x, w = gauss_legendre(n)
    irv = np.zeros(r_steps)
    rp = np.zeros((r_steps, n))
    rm = np.zeros((r_steps, n))
    rm2 = np.zeros((r_steps, n))
    qw = np.zeros((r_steps, n))
    reff = np.zeros(r_steps)
    nus = np.zeros(nu_steps)
    ids = np.zeros(nu_steps)
    lib().orc_r_constants(n, _ptr(x), _ptr(w), r_steps, r_min, r_step, mu_step, nu_steps,
                          nu_min, nu_step, _ptr(irv), _ptr(rp), _ptr(rm), _ptr(rm2), _ptr(qw),
                          _ptr(reff), _ptr(nus), _ptr(ids))
    return dict(irv=irv, r_point=rp, r_in_mag=rm, r_in_mag2=rm2, qw_r3_N=qw,
                reff_xr_rp3=reff, nus=nus, ids=ids)


def point_probabilities(params: ParamsHolder, l: float, b: float, r_kpc: float):
    S = params.ap.number_streams
    bg = C.c_double()
    st = np.zeros(max(1, S))
    rc = lib().orc_point_probabilities(params.ptr, l, b, r_kpc, C.byref(bg), _ptr(st))
    if rc:
        raise ValueError("invalid parameters")
    return bg.value, st[:S]


class OracleResult:
    def __init__(self, state: StateHolder, n_stars: int, log10_p=None, bg_only=None,
                 st_only=None):
        es = state.es
        self.background_integral = es.background_integral
        self.stream_integrals = state.stream_integrals()
        self.regions = [state.region_integrals(i) for i in range(es.number_integrals)]
        self.prob_sum = es.prob_sum
        self.num_zero = es.num_zero
        self.likelihood = es.prob_sum / (n_stars - es.bad_jacobians) if n_stars else float("nan")
        self.log10_p = log10_p
        self.bg_only_mean = bg_only
        self.st_only_mean = st_only


def integrals(params: ParamsHolder, stars: StarsHolder, state: StateHolder = None,
              threads: int = 0) -> StateHolder:
    state = state or StateHolder(params, stars)
    if threads:
        rc = lib().orc_calculate_integrals_omp(params.ptr, state.ptr, threads)
    else:
        rc = lib().orc_calculate_integrals(params.ptr, state.ptr, stars.ptr)
    if rc:
        raise ValueError("oracle rejected the parameters")
    return state


def likelihood(params: ParamsHolder, stars: StarsHolder, state: StateHolder,
               threads: int = 0) -> OracleResult:
    n = len(stars)
    S = params.ap.number_streams
    if threads:
        rc = lib().orc_calculate_likelihood_omp(params.ptr, state.ptr, stars.ptr, threads)
        if rc:
            raise ValueError("oracle rejected the parameters")
        return OracleResult(state, n)
    lp = np.zeros(max(1, n))
    bgm = C.c_double()
    stm = np.zeros(max(1, S))
    rc = lib().orc_calculate_likelihood_ex(params.ptr, state.ptr, stars.ptr, _ptr(lp),
                                           C.byref(bgm), _ptr(stm))
    if rc:
        raise ValueError("oracle rejected the parameters")
    return OracleResult(state, n, lp[:n], bgm.value, stm[:S])


def evaluate_full(params: ParamsHolder, stars: StarsHolder, vector=None,
                  threads: int = 0) -> OracleResult:
    """cpu_evaluate semantics (evaluation_optimized.c:756-784) with per-star
    outputs: optional parameter vector, integrals, then the star pass."""
    if vector is not None:
        v = np.ascontiguousarray(vector, dtype=np.float64)
        lib().orc_set_astronomy_parameters(params.ptr, _ptr(v))
    state = StateHolder(params, stars)
    integrals(params, stars, state, threads)
    return likelihood(params, stars, state, threads)


def evaluate(params: ParamsHolder, stars: StarsHolder, vector) -> float:
    """orc_evaluate: the reference evaluator's return value (no -3.0)."""
    state = StateHolder(params, stars)
    v = np.ascontiguousarray(vector, dtype=np.float64)
    return lib().orc_evaluate(_ptr(v), params.ptr, state.ptr, stars.ptr)


def omp_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def wedge_incl(wedge: int) -> float:
    return lib().orc_wedge_incl(wedge)


def get_node() -> float:
    return lib().orc_get_node()


def stripe_normal(wedge: int, sgr: int = 0) -> np.ndarray:
    """stCoords.c:142-174: the wedge's normal (SDSS or Sagittarius stripe)."""
    v = np.zeros(3)
    (lib().orc_sgr_stripe_normal if sgr else lib().orc_stripe_normal)(wedge, _ptr(v))
    return v


def get_transform(f, t) -> np.ndarray:
    """stVector.c:109-140: rotation matrix taking f to t."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    m = np.zeros(9)
    lib().orc_get_transform(_ptr(f), _ptr(t), _ptr(m))
    return m.reshape(3, 3)


def transform_point(point, cmat, xsun=(-8.5, 0.0, 0.0)) -> np.ndarray:
    """stVector.c:155-169 on a galactocentric point."""
    p = np.ascontiguousarray(point, dtype=np.float64)
    m = np.ascontiguousarray(cmat, dtype=np.float64).reshape(9)
    xs = np.ascontiguousarray(xsun, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_transform_point(_ptr(p), _ptr(m), _ptr(xs), _ptr(out))
    return out


def prob_ok(p, r: float) -> int:
    """probability.c:310-396 with the random number r given (-1 where the
    reference leaves the result unset)."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    return lib().orc_prob_ok(len(p), _ptr(p), r)


class SeparationResult:
    def __init__(self, assign, xyz, sprob, nstars, q):
        self.assignment = assign
        self.xyz = xyz
        self.probability = sprob
        self.expected = nstars
        self.separated = q


def separation(params: ParamsHolder, stars: StarsHolder, background_integral: float,
               stream_integrals, seed: int) -> SeparationResult:
    """separation.c:37-254 for the stars with the given integrals and
    srand48(seed)."""
    S = params.ap.number_streams
    N = len(stars)
    si = np.ascontiguousarray(stream_integrals, dtype=np.float64)
    assign = np.zeros(max(1, N), dtype=np.int32)
    xyz = np.zeros((max(1, N), 3))
    sprob = np.zeros((max(1, N), max(1, S)))
    nstars = np.zeros(max(1, S))
    q = np.zeros(max(1, S), dtype=np.int32)
    ip = C.POINTER(C.c_int)
    rc = lib().orc_separation(params.ptr, stars.ptr, background_integral, _ptr(si), seed,
                              assign.ctypes.data_as(ip), _ptr(xyz), _ptr(sprob), _ptr(nstars),
                              q.ctypes.data_as(ip))
    if rc:
        raise ValueError("oracle rejected the separation inputs")
    return SeparationResult(assign[:N], xyz[:N], sprob[:N, :S], nstars[:S], q[:S])
