"""Python host interface of the HIP separation library.

Mirrors the reference evaluator's interface
(``milkyway_separation/cpu/evaluation_optimized.h:49-62``) over the C ABI of
``include/mw_hip.h``:

* ``evaluate``              -> ``mw_hip_evaluate``   (cpu_evaluate)
* ``calculate_integrals``   -> ``mw_hip_calculate_integrals``
* ``calculate_likelihood``  -> ``mw_hip_calculate_likelihood``
* ``Context``               -> resident wedge + batched evaluation

The library is the only compute path: if it is not built, or no GPU is
visible, these calls raise instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

from .refstructs import (ASTRONOMY_PARAMETERS, EVALUATION_STATE, STAR_POINTS, ParamsHolder,
                         StarsHolder, StateHolder)

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libmw_separation_hip.so"

MW_OK, MW_E_INVALID, MW_E_HIP, MW_E_NOMEM, MW_E_NODEVICE = 0, 1, 2, 3, 4
PREC_FP64, PREC_MIXED = 0, 1
MAX_STREAMS = 4
CTX_TIMING = 1

# Every symbol the C ABI exports (include/mw_hip.h).
EXPORTS = [
    "mw_hip_evaluate", "mw_hip_calculate_integrals", "mw_hip_calculate_likelihood",
    "mw_hip_set_precision", "mw_ctx_create", "mw_ctx_parameter_count", "mw_ctx_eval",
    "mw_ctx_region_integrals", "mw_ctx_last_timing", "mw_ctx_resident_bytes",
    "mw_ctx_destroy", "mw_last_error", "mw_device_count", "mw_device_synchronize",
]


class MwError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mw error {code}: {msg}")
        self.code = code


class mw_result(C.Structure):
    _fields_ = [
        ("likelihood", C.c_double), ("prob_sum", C.c_double), ("num_zero", C.c_int),
        ("number_streams", C.c_int), ("background_integral", C.c_double),
        ("stream_integrals", C.c_double * MAX_STREAMS),
        ("bg_only_likelihood", C.c_double),
        ("stream_only_likelihood", C.c_double * MAX_STREAMS),
    ]


class mw_timing(C.Structure):
    _fields_ = [("integral_ms", C.c_double), ("star_ms", C.c_double), ("total_ms", C.c_double),
                ("integral_blocks", C.c_long), ("star_blocks", C.c_long)]


_lib = None
_dp = C.POINTER(C.c_double)


def load_library() -> C.CDLL:
    """Load the built library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"HIP library not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(str(LIB_PATH))
    AP, ES, SP = (C.POINTER(ASTRONOMY_PARAMETERS), C.POINTER(EVALUATION_STATE),
                  C.POINTER(STAR_POINTS))
    L.mw_hip_evaluate.argtypes = [_dp, AP, ES, SP]
    L.mw_hip_evaluate.restype = C.c_double
    L.mw_hip_calculate_integrals.argtypes = [AP, ES, SP]
    L.mw_hip_calculate_integrals.restype = C.c_int
    L.mw_hip_calculate_likelihood.argtypes = [AP, ES, SP]
    L.mw_hip_calculate_likelihood.restype = C.c_int
    L.mw_hip_set_precision.argtypes = [C.c_int]
    L.mw_hip_set_precision.restype = C.c_int
    L.mw_ctx_create.argtypes = [AP, SP, C.c_int, C.POINTER(C.c_void_p)]
    L.mw_ctx_create.restype = C.c_int
    L.mw_ctx_parameter_count.argtypes = [C.c_void_p]
    L.mw_ctx_parameter_count.restype = C.c_int
    L.mw_ctx_eval.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.POINTER(mw_result), _dp]
    L.mw_ctx_eval.restype = C.c_int
    L.mw_ctx_region_integrals.argtypes = [C.c_void_p, _dp, C.c_size_t]
    L.mw_ctx_region_integrals.restype = C.c_int
    L.mw_ctx_last_timing.argtypes = [C.c_void_p, C.POINTER(mw_timing)]
    L.mw_ctx_last_timing.restype = C.c_int
    L.mw_ctx_resident_bytes.argtypes = [C.c_void_p]
    L.mw_ctx_resident_bytes.restype = C.c_size_t
    L.mw_ctx_destroy.argtypes = [C.c_void_p]
    L.mw_ctx_destroy.restype = None
    L.mw_last_error.argtypes = []
    L.mw_last_error.restype = C.c_char_p
    L.mw_device_count.argtypes = []
    L.mw_device_count.restype = C.c_int
    L.mw_device_synchronize.argtypes = []
    L.mw_device_synchronize.restype = C.c_int
    _lib = L
    return L


def _check(rc: int):
    if rc != MW_OK:
        msg = load_library().mw_last_error().decode(errors="replace")
        raise MwError(rc, msg)


def device_count() -> int:
    return load_library().mw_device_count()


def require_device():
    if device_count() <= 0:
        raise RuntimeError("no GPU visible: the HIP separation library needs one")


def synchronize():
    _check(load_library().mw_device_synchronize())


def set_precision(precision: int):
    _check(load_library().mw_hip_set_precision(precision))


def evaluate(params: ParamsHolder, state: StateHolder, stars: StarsHolder, vector) -> float:
    """cpu_evaluate replacement: returns prob_sum / number_stars (no -3.0)."""
    v = np.ascontiguousarray(vector, dtype=np.float64)
    return load_library().mw_hip_evaluate(v.ctypes.data_as(_dp), params.ptr, state.ptr, stars.ptr)


def calculate_integrals(params: ParamsHolder, state: StateHolder, stars: Optional[StarsHolder]):
    _check(load_library().mw_hip_calculate_integrals(params.ptr, state.ptr,
                                                     stars.ptr if stars else None))


def calculate_likelihood(params: ParamsHolder, state: StateHolder, stars: StarsHolder):
    _check(load_library().mw_hip_calculate_likelihood(params.ptr, state.ptr, stars.ptr))


class Result:
    """One evaluation's outputs (mw_result), plus per-star log10 P if asked."""

    def __init__(self, r: mw_result, log10p=None, regions=None):
        S = r.number_streams
        self.likelihood = r.likelihood
        self.prob_sum = r.prob_sum
        self.num_zero = r.num_zero
        self.background_integral = r.background_integral
        self.stream_integrals = np.array(r.stream_integrals[:S])
        self.bg_only_likelihood = r.bg_only_likelihood
        self.stream_only_likelihood = np.array(r.stream_only_likelihood[:S])
        self.log10_p = log10p
        self.regions = regions


class Context:
    """A wedge (regions, tables) and star list resident on the current GPU."""

    def __init__(self, params: ParamsHolder, stars: Optional[StarsHolder], timing: bool = False):
        L = load_library()
        self._params = params
        self._stars = stars
        h = C.c_void_p()
        _check(L.mw_ctx_create(params.ptr, stars.ptr if stars else None,
                               CTX_TIMING if timing else 0, C.byref(h)))
        self.handle = h
        self.n_stars = len(stars) if stars else 0
        self.n_regions = params.ap.number_integrals
        self.n_params = L.mw_ctx_parameter_count(h)

    def close(self):
        if self.handle:
            load_library().mw_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def resident_bytes(self) -> int:
        return load_library().mw_ctx_resident_bytes(self.handle)

    def eval(self, vectors, precision: int = PREC_FP64, per_star: bool = False,
             regions: bool = False) -> List[Result]:
        vecs = np.ascontiguousarray(np.atleast_2d(np.asarray(vectors, dtype=np.float64)))
        if self.n_params == 0:
            vecs = np.zeros((max(1, vecs.shape[0]), 1))
        elif vecs.shape[1] != self.n_params:
            raise ValueError(f"parameter vectors need {self.n_params} values, got {vecs.shape[1]}")
        n = vecs.shape[0]
        res = (mw_result * n)()
        lp = np.zeros((n, max(1, self.n_stars))) if per_star else None
        _check(load_library().mw_ctx_eval(self.handle, vecs.ctypes.data_as(_dp), n, precision,
                                          res, lp.ctypes.data_as(_dp) if per_star else None))
        reg = None
        if regions:
            reg = np.zeros((n, self.n_regions, 1 + MAX_STREAMS))
            _check(load_library().mw_ctx_region_integrals(self.handle, reg.ctypes.data_as(_dp),
                                                          reg.size))
        return [Result(res[i], lp[i, :self.n_stars] if per_star else None,
                       reg[i] if regions else None) for i in range(n)]

    def timing(self) -> mw_timing:
        t = mw_timing()
        _check(load_library().mw_ctx_last_timing(self.handle, C.byref(t)))
        return t
