"""mw_ctx_eval_noalloc: one evaluation that allocates nothing, its working
memory a fixed device array.  It runs the same kernels as mw_ctx_eval for
one vector, so its result must be mw_ctx_eval's bitwise (the FP64 parity
bar, integrals 1e-12 and the likelihood 1e-10 against the oracle, follows
from that).  Checked on two calls in a row through the C ABI, the second on
a smaller wedge with fewer stars, so that it reuses what the first left in
the array; and sizes beyond the array are refused.
"""
import ctypes as C

import numpy as np
import pytest

from milkyway_client_amd import separation as sep
from milkyway_client_amd import synth
from milkyway_client_amd.refstructs import ParamsHolder, StarsHolder

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def noalloc(ctx, vec, precision):
    L = sep.load_library()
    L.mw_ctx_eval_noalloc.argtypes = [C.c_void_p, _dp, C.c_int, C.c_void_p, C.POINTER(sep.mw_result)]
    L.mw_ctx_eval_noalloc.restype = C.c_int
    v = np.ascontiguousarray(vec, dtype=np.float64)
    r = sep.mw_result()
    rc = L.mw_ctx_eval_noalloc(ctx.handle, v.ctypes.data_as(_dp), precision, None, C.byref(r))
    if rc:
        raise sep.MwError(rc, L.mw_last_error().decode())
    return sep.Result(r)


def same(a, b):
    return (a.likelihood == b.likelihood and a.prob_sum == b.prob_sum and a.num_zero == b.num_zero
            and a.background_integral == b.background_integral
            and np.array_equal(a.stream_integrals, b.stream_integrals)
            and a.bg_only_likelihood == b.bg_only_likelihood
            and np.array_equal(a.stream_only_likelihood, b.stream_only_likelihood))


@pytest.mark.parametrize("precision", [sep.PREC_FP64, sep.PREC_MIXED])
def test_noalloc_equals_ctx_eval_on_two_calls(precision):
    sep.load_library()
    sep.require_device()
    big = synth.workload(300, synth.WedgeShape(70, 160, 64, 30, 10_000,
                                                cut=(70, 40, 64, 330.0, 355.0)), 3)
    small = synth.workload(301, synth.WedgeShape(20, 24, 16, 15, 700), 2)
    for p, stars in (big, small):
        with sep.Context(ParamsHolder(p), StarsHolder(stars)) as ctx:
            vec = synth.parameter_vectors(302, p, 1)[0]
            want = ctx.eval(vec, precision=precision)[0]
            got = noalloc(ctx, vec, precision)
            assert same(got, want)


def test_noalloc_refuses_what_it_cannot_hold():
    sep.load_library()
    sep.require_device()
    # the full-size wedge needs about 0.59 MB of scratch and fits; a model of
    # more than 4 streams and the resampled quadrature are refused
    p, stars = synth.workload(303, synth.WedgeShape(10, 12, 8, 10, 50), 5)
    with sep.Context(ParamsHolder(p), StarsHolder(stars)) as ctx:
        with pytest.raises(sep.MwError):
            noalloc(ctx, p.parameter_vector(), sep.PREC_FP64)
    p, stars = synth.workload(304, synth.WedgeShape(10, 12, 8, 10, 50), 2)
    with sep.Context(ParamsHolder(p), StarsHolder(stars)) as ctx:
        with pytest.raises(sep.MwError):
            noalloc(ctx, p.parameter_vector(), sep.PREC_RESAMPLED)
    # a wedge of 14 M sight columns needs more than the 4 MiB array
    p, _ = synth.workload(305, synth.WedgeShape(2, 7000, 2048, 1, 0), 1)
    with sep.Context(ParamsHolder(p), None) as ctx:
        with pytest.raises(sep.MwError, match="scratch"):
            noalloc(ctx, p.parameter_vector(), sep.PREC_FP64)
