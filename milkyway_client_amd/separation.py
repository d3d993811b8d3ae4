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
# The command-line client (src/milkyway_client_app.c's worker on this library).
CLIENT_PATH = LIB_PATH.parent / "milkyway_separation_hip"

MW_OK, MW_E_INVALID, MW_E_HIP, MW_E_NOMEM, MW_E_NODEVICE = 0, 1, 2, 3, 4
PREC_FP64, PREC_MIXED, PREC_RESAMPLED = 0, 1, 2
MAX_STREAMS = 16   # MW_MAX_STREAMS: streams per model in the context API
CTX_TIMING = 1

# Every symbol the C ABI exports (include/mw_hip.h).
EXPORTS = [
    "mw_hip_evaluate", "mw_hip_calculate_integrals", "mw_hip_calculate_likelihood",
    "mw_hip_set_precision", "mw_ctx_create", "mw_ctx_parameter_count", "mw_ctx_eval",
    "mw_ctx_region_integrals", "mw_ctx_last_timing", "mw_ctx_resident_bytes",
    "mw_ctx_destroy", "mw_last_error", "mw_device_count", "mw_set_device", "mw_device_synchronize",
    "mw_batch_create", "mw_batch_run", "mw_batch_set_variant", "mw_batch_variant_count",
    "mw_batch_variant_scratch",
    "mw_batch_vector_count", "mw_batch_last_timing", "mw_batch_destroy",
    "mw_server_create", "mw_server_add_wedge", "mw_server_submit", "mw_server_poll",
    "mw_server_wait", "mw_server_get_stats", "mw_server_destroy",
    "mw_read_astronomy_parameters", "mw_free_astronomy_parameters", "mw_read_star_points",
    "mw_free_star_points", "mw_initialize_state", "mw_free_state",
    "mw_get_optimized_parameter_count", "mw_set_astronomy_parameters",
    "mw_write_astronomy_parameters", "mw_write_star_points",
    "mw_ctx_separation", "mw_hip_separation",
    "mw_hip_integral_f", "mw_hip_likelihood_f", "mw_integral_compose", "mw_likelihood_compose",
    "mw_split_astronomy_parameters", "mw_split_star_points", "mw_get_search_parameters",
    "mw_get_step", "mw_get_min_parameters", "mw_get_max_parameters", "mw_hessian",
    "mw_ctx_hessian", "mw_matrix_invert", "mw_errors",
    "mw_write_checkpoint", "mw_read_checkpoint", "mw_calculate_progress",
    "mw_resample_weights", "mw_resample_weights_padded", "mw_ctx_eval_noalloc",
]

# int (*)(void* user, const double* vectors, int count, double* values)
VECTOR_EVAL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int,
                             C.POINTER(C.c_double))


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
                ("integral_blocks", C.c_long), ("star_blocks", C.c_long),
                ("integral_points", C.c_long), ("integral_points_recurrence", C.c_long),
                ("integral_points_background", C.c_long), ("integral_stream_terms", C.c_long)]


class mw_server_config(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_inflight", C.c_int),
                ("max_batch_vectors", C.c_int), ("variant", C.c_int), ("blocks_per_cu", C.c_int)]


class mw_separation_summary(C.Structure):
    _fields_ = [("number_streams", C.c_int), ("number_stars", C.c_int),
                ("epsilon_b", C.c_double), ("epsilon_s", C.c_double * MAX_STREAMS),
                ("expected", C.c_double * MAX_STREAMS), ("separated", C.c_int * MAX_STREAMS),
                ("background_integral", C.c_double),
                ("stream_integrals", C.c_double * MAX_STREAMS)]


class mw_request_times(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("dispatch_s", C.c_double), ("done_s", C.c_double)]


class mw_server_stats(C.Structure):
    _fields_ = [("resident_bytes", C.c_size_t), ("inflight_scratch_bytes", C.c_size_t),
                ("pending", C.c_long), ("batches", C.c_long), ("completed", C.c_long),
                ("largest_batch", C.c_long), ("inflight_batches", C.c_int)]


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
    L.mw_set_device.argtypes = [C.c_int]
    L.mw_set_device.restype = C.c_int
    L.mw_device_synchronize.argtypes = []
    L.mw_device_synchronize.restype = C.c_int
    L.mw_batch_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.c_int, C.c_int,
                                  C.c_int, C.POINTER(C.c_void_p)]
    L.mw_batch_create.restype = C.c_int
    L.mw_batch_run.argtypes = [C.c_void_p, _dp, C.POINTER(mw_result)]
    L.mw_batch_run.restype = C.c_int
    L.mw_batch_set_variant.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.mw_batch_set_variant.restype = C.c_int
    L.mw_batch_variant_count.argtypes = []
    L.mw_batch_variant_count.restype = C.c_int
    L.mw_batch_variant_scratch.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_size_t)]
    L.mw_batch_variant_scratch.restype = C.c_int
    L.mw_batch_vector_count.argtypes = [C.c_void_p]
    L.mw_batch_vector_count.restype = C.c_int
    L.mw_batch_last_timing.argtypes = [C.c_void_p, C.POINTER(mw_timing)]
    L.mw_batch_last_timing.restype = C.c_int
    L.mw_batch_destroy.argtypes = [C.c_void_p]
    L.mw_batch_destroy.restype = None
    AP, SP = C.POINTER(ASTRONOMY_PARAMETERS), C.POINTER(STAR_POINTS)
    L.mw_server_create.argtypes = [C.POINTER(mw_server_config), C.POINTER(C.c_void_p)]
    L.mw_server_create.restype = C.c_int
    L.mw_server_add_wedge.argtypes = [C.c_void_p, AP, SP, C.POINTER(C.c_int)]
    L.mw_server_add_wedge.restype = C.c_int
    L.mw_server_submit.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, C.POINTER(C.c_long)]
    L.mw_server_submit.restype = C.c_int
    L.mw_server_poll.argtypes = [C.c_void_p, C.c_long, C.POINTER(mw_result),
                                 C.POINTER(mw_request_times), C.POINTER(C.c_int)]
    L.mw_server_poll.restype = C.c_int
    L.mw_server_wait.argtypes = [C.c_void_p, C.c_long, C.POINTER(mw_result),
                                 C.POINTER(mw_request_times)]
    L.mw_server_wait.restype = C.c_int
    L.mw_server_get_stats.argtypes = [C.c_void_p, C.POINTER(mw_server_stats)]
    L.mw_server_get_stats.restype = C.c_int
    L.mw_server_destroy.argtypes = [C.c_void_p]
    L.mw_server_destroy.restype = None
    ES = C.POINTER(EVALUATION_STATE)
    L.mw_read_astronomy_parameters.argtypes = [C.c_char_p, AP]
    L.mw_read_astronomy_parameters.restype = C.c_int
    L.mw_free_astronomy_parameters.argtypes = [AP]
    L.mw_free_astronomy_parameters.restype = None
    L.mw_read_star_points.argtypes = [C.c_char_p, SP]
    L.mw_read_star_points.restype = C.c_int
    L.mw_free_star_points.argtypes = [SP]
    L.mw_free_star_points.restype = None
    L.mw_initialize_state.argtypes = [AP, SP, ES]
    L.mw_initialize_state.restype = C.c_int
    L.mw_free_state.argtypes = [ES]
    L.mw_free_state.restype = None
    L.mw_get_optimized_parameter_count.argtypes = [AP]
    L.mw_get_optimized_parameter_count.restype = C.c_int
    L.mw_set_astronomy_parameters.argtypes = [AP, _dp]
    L.mw_set_astronomy_parameters.restype = None
    L.mw_write_astronomy_parameters.argtypes = [C.c_char_p, AP]
    L.mw_write_astronomy_parameters.restype = C.c_int
    L.mw_write_star_points.argtypes = [C.c_char_p, SP]
    L.mw_write_star_points.restype = C.c_int
    L.mw_ctx_separation.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_ulong, C.POINTER(C.c_int),
                                    _dp, _dp, C.POINTER(mw_separation_summary)]
    L.mw_ctx_separation.restype = C.c_int
    L.mw_hip_separation.argtypes = [C.c_char_p, C.c_double, _dp, AP, SP, C.c_ulong]
    L.mw_hip_separation.restype = C.c_int
    L.mw_hip_integral_f.argtypes = [_dp, _dp, AP, ES, SP]
    L.mw_hip_integral_f.restype = None
    L.mw_hip_likelihood_f.argtypes = [_dp, _dp, AP, ES, SP]
    L.mw_hip_likelihood_f.restype = None
    L.mw_integral_compose.argtypes = [_dp, C.c_int, _dp, AP]
    L.mw_integral_compose.restype = None
    L.mw_likelihood_compose.argtypes = [_dp, C.c_size_t, C.c_int]
    L.mw_likelihood_compose.restype = C.c_double
    L.mw_split_astronomy_parameters.argtypes = [AP, C.c_int, C.c_int]
    L.mw_split_astronomy_parameters.restype = None
    L.mw_split_star_points.argtypes = [SP, C.c_int, C.c_int]
    L.mw_split_star_points.restype = C.c_int
    for name in ("mw_get_search_parameters", "mw_get_step", "mw_get_min_parameters",
                 "mw_get_max_parameters"):
        getattr(L, name).argtypes = [AP, _dp]
        getattr(L, name).restype = None
    L.mw_hessian.argtypes = [C.c_int, _dp, _dp, VECTOR_EVAL_FN, C.c_void_p, _dp, _dp, _dp]
    L.mw_hessian.restype = C.c_int
    L.mw_ctx_hessian.argtypes = [C.c_void_p, _dp, _dp, C.c_int, _dp, _dp, _dp]
    L.mw_ctx_hessian.restype = C.c_int
    L.mw_matrix_invert.argtypes = [C.c_int, _dp, _dp]
    L.mw_matrix_invert.restype = C.c_int
    L.mw_errors.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp]
    L.mw_errors.restype = C.c_int
    L.mw_write_checkpoint.argtypes = [C.c_char_p, ES]
    L.mw_write_checkpoint.restype = C.c_int
    L.mw_read_checkpoint.argtypes = [C.c_char_p, ES]
    L.mw_read_checkpoint.restype = C.c_int
    L.mw_calculate_progress.argtypes = [ES]
    L.mw_calculate_progress.restype = C.c_double
    L.mw_resample_weights.argtypes = [_dp, _dp, C.c_long, C.c_double, C.c_int, _dp, _dp, C.c_long]
    L.mw_resample_weights.restype = C.c_long
    L.mw_resample_weights_padded.argtypes = [_dp, _dp, C.c_long, C.c_double, C.c_int, C.c_int, _dp, _dp,
                                             C.c_long]
    L.mw_resample_weights_padded.restype = C.c_long
    _lib = L
    return L


def _check(rc: int):
    if rc != MW_OK:
        msg = load_library().mw_last_error().decode(errors="replace")
        raise MwError(rc, msg)


def device_count() -> int:
    return load_library().mw_device_count()


def set_device(device: int):
    _check(load_library().mw_set_device(device))


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


def hip_separation(filename, background_integral: float, stream_integrals,
                   params: ParamsHolder, stars: StarsHolder, seed: int) -> None:
    """separation() replacement (mw_hip_separation): writes the output file."""
    si = np.ascontiguousarray(stream_integrals, dtype=np.float64)
    _check(load_library().mw_hip_separation(str(filename).encode(), background_integral,
                                            si.ctypes.data_as(_dp), params.ptr, stars.ptr, seed))


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

    def hessian(self, point, step, precision: int = PREC_FP64):
        """Central-difference Hessian, gradient and value of the client's
        likelihood (mw_ctx_hessian), all 2 n^2 + 1 points in batched launches."""
        n = self.n_params
        x = np.ascontiguousarray(point, dtype=np.float64)
        h = np.ascontiguousarray(step, dtype=np.float64)
        if x.size != n or h.size != n:
            raise ValueError(f"point and step need {n} values")
        H = np.zeros((n, n))
        g = np.zeros(n)
        v = C.c_double()
        _check(load_library().mw_ctx_hessian(self.handle, x.ctypes.data_as(_dp), h.ctypes.data_as(_dp),
                                             precision, H.ctypes.data_as(_dp), g.ctypes.data_as(_dp),
                                             C.byref(v)))
        return H, g, v.value

    def separation(self, vector=None, precision: int = PREC_FP64, integrals=None, seed: int = 0):
        """Stream separation (mw_ctx_separation): per-star assignment [N],
        position [N][3] and membership probabilities [N][S], and the summary."""
        S = self._params.ap.number_streams
        N = self.n_stars
        vec = None
        if vector is not None:
            vec = np.ascontiguousarray(vector, dtype=np.float64)
            if vec.size != self.n_params:
                raise ValueError(f"parameter vector needs {self.n_params} values, got {vec.size}")
        ints = None
        if integrals is not None:
            ints = np.ascontiguousarray(integrals, dtype=np.float64)
            if ints.size != 1 + S:
                raise ValueError(f"integrals need {1 + S} values")
        assign = np.zeros(max(1, N), dtype=np.int32)
        xyz = np.zeros((max(1, N), 3))
        prob = np.zeros((max(1, N), max(1, S)))
        summ = mw_separation_summary()
        _check(load_library().mw_ctx_separation(
            self.handle, vec.ctypes.data_as(_dp) if vec is not None else None, precision,
            ints.ctypes.data_as(_dp) if ints is not None else None, seed,
            assign.ctypes.data_as(C.POINTER(C.c_int)), xyz.ctypes.data_as(_dp),
            prob.ctypes.data_as(_dp), C.byref(summ)))
        return assign[:N], xyz[:N], prob[:N, :S], summ


BATCH_GRAPH = 1


class Batch:
    """Parameter vectors over several resident wedges in one persistent launch
    (mw_batch_*).  ``groups`` is a list of (Context, n_vectors)."""

    def __init__(self, groups, precision: int = PREC_MIXED, graph: bool = True):
        L = load_library()
        self.contexts = [c for c, _ in groups]
        self.counts = [int(n) for _, n in groups]
        arr = (C.c_void_p * len(groups))(*[c.handle.value for c in self.contexts])
        cnt = (C.c_int * len(groups))(*self.counts)
        h = C.c_void_p()
        _check(L.mw_batch_create(arr, cnt, len(groups), precision,
                                 BATCH_GRAPH if graph else 0, C.byref(h)))
        self.handle = h
        self.n_vectors = sum(self.counts)

    @staticmethod
    def variant_count() -> int:
        return load_library().mw_batch_variant_count()

    def set_variant(self, variant: int, blocks_per_cu: int):
        _check(load_library().mw_batch_set_variant(self.handle, variant, blocks_per_cu))

    def variant_scratch(self, variant: int) -> int:
        """Scratch (register spill) bytes per work-item of this batch's
        integral kernels at the given variant."""
        n = C.c_size_t()
        _check(load_library().mw_batch_variant_scratch(self.handle, variant, C.byref(n)))
        return int(n.value)

    def run(self, vectors) -> List[Result]:
        """vectors: one array per context, (n_vectors, n_params) each."""
        parts = []
        for ctx, n, v in zip(self.contexts, self.counts, vectors):
            a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(n, max(ctx.n_params, 0)))
            parts.append(a.ravel())
        flat = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(1))
        res = (mw_result * self.n_vectors)()
        _check(load_library().mw_batch_run(self.handle, flat.ctypes.data_as(_dp), res))
        return [Result(res[i]) for i in range(self.n_vectors)]

    def timing(self) -> mw_timing:
        t = mw_timing()
        _check(load_library().mw_batch_last_timing(self.handle, C.byref(t)))
        return t

    def close(self):
        if self.handle:
            load_library().mw_batch_destroy(self.handle)
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


class Server:
    """Request server (mw_server_*): resident wedges, queued requests admitted
    into concurrent batches by an HBM budget."""

    def __init__(self, budget_bytes: int = 0, max_inflight: int = 3, max_batch: int = 64,
                 variant: int = 0, blocks_per_cu: int = 2):
        cfg = mw_server_config(budget_bytes, max_inflight, max_batch, variant, blocks_per_cu)
        h = C.c_void_p()
        _check(load_library().mw_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h
        self._keep = []
        self.n_params = []

    def add_wedge(self, params: ParamsHolder, stars: Optional[StarsHolder]) -> int:
        wid = C.c_int()
        _check(load_library().mw_server_add_wedge(self.handle, params.ptr,
                                                  stars.ptr if stars else None, C.byref(wid)))
        self._keep.append((params, stars))
        self.n_params.append(params.py.optimized_parameter_count())
        return wid.value

    def submit(self, wedge_id: int, vector, precision: int = PREC_MIXED) -> int:
        v = np.ascontiguousarray(np.asarray(vector, dtype=np.float64).ravel())
        if v.size == 0:
            v = np.zeros(1)
        rid = C.c_long()
        _check(load_library().mw_server_submit(self.handle, wedge_id, v.ctypes.data_as(_dp),
                                               precision, C.byref(rid)))
        return rid.value

    def poll(self, request_id: int):
        r, t, done = mw_result(), mw_request_times(), C.c_int()
        _check(load_library().mw_server_poll(self.handle, request_id, C.byref(r), C.byref(t),
                                             C.byref(done)))
        if not done.value:
            return None
        return Result(r), (t.submit_s, t.dispatch_s, t.done_s)

    def wait(self, request_id: int):
        r, t = mw_result(), mw_request_times()
        _check(load_library().mw_server_wait(self.handle, request_id, C.byref(r), C.byref(t)))
        return Result(r), (t.submit_s, t.dispatch_s, t.done_s)

    def stats(self) -> mw_server_stats:
        st = mw_server_stats()
        _check(load_library().mw_server_get_stats(self.handle, C.byref(st)))
        return st

    def close(self):
        if self.handle:
            load_library().mw_server_destroy(self.handle)
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


def errors(hessian, number_stars: int):
    """errors.c:38-84: sigma = H^-1 / number_stars and the standard deviations."""
    H = np.ascontiguousarray(hessian, dtype=np.float64)
    n = H.shape[0]
    sigma = np.zeros((n, n))
    sd = np.zeros(n)
    if load_library().mw_errors(n, H.ctypes.data_as(_dp), number_stars, sigma.ctypes.data_as(_dp),
                                sd.ctypes.data_as(_dp)):
        raise np.linalg.LinAlgError("singular Hessian")
    return sigma, sd


def resample_weights(g, w, delta: float, degree: int = 15, pad: int = None):
    """mw_resample_weights (pad None: every stencil centred) or
    mw_resample_weights_padded: grid start, grid weights K."""
    g = np.ascontiguousarray(g, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    lo = C.c_double()
    L = load_library()
    if pad is None:
        pad = degree // 2 + 1
    M = L.mw_resample_weights_padded(g.ctypes.data_as(_dp), w.ctypes.data_as(_dp), g.size, delta, degree,
                                     pad, C.byref(lo), None, 0)
    if M < 1:
        raise ValueError("bad resampling input")
    K = np.zeros(M)
    L.mw_resample_weights_padded(g.ctypes.data_as(_dp), w.ctypes.data_as(_dp), g.size, delta, degree,
                                 pad, C.byref(lo), K.ctypes.data_as(_dp), M)
    return lo.value, K
