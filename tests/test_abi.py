"""The C ABI: the library loads without a GPU and exports every function that
include/*.h declares, and the public structs are byte-identical to the
reference's (milkyway_separation/cpu/parameters.h, evaluation_state.h,
star_points.h) -- checked by compiling offsetof tables against both headers
when the reference tree is mounted."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from milkyway_client_amd import refstructs
from milkyway_client_amd import separation as sep

REPO = Path(__file__).resolve().parents[1]
REFERENCE = Path("/root/reference/milkyway_separation/cpu")


def declared_functions():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(mw_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    lib = C.CDLL(str(sep.LIB_PATH))
    names = declared_functions()
    assert names, "no declarations found"
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(sep.EXPORTS)


def test_library_reports_no_device_gracefully():
    n = sep.device_count()
    assert n >= 0
    if n == 0:
        with pytest.raises(RuntimeError):
            sep.require_device()


def test_ctypes_mirror_sizes():
    assert C.sizeof(refstructs.INTEGRAL) == 104
    assert C.sizeof(refstructs.STAR_POINTS) == 16
    assert C.sizeof(refstructs.ASTRONOMY_PARAMETERS) == 184


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
%(include)s
#define F(T, f) printf(#T "." #f " %%zu\n", offsetof(T, f));
#define Z(T) printf(#T " %%zu\n", sizeof(T));
int main(void) {
  Z(INTEGRAL) F(INTEGRAL, r_min) F(INTEGRAL, nu_step_size) F(INTEGRAL, r_steps) F(INTEGRAL, mu_steps)
  F(INTEGRAL, min_calculation) F(INTEGRAL, max_calculation)
  Z(ASTRONOMY_PARAMETERS) F(ASTRONOMY_PARAMETERS, number_background_parameters)
  F(ASTRONOMY_PARAMETERS, background_weight) F(ASTRONOMY_PARAMETERS, background_parameters)
  F(ASTRONOMY_PARAMETERS, background_optimize) F(ASTRONOMY_PARAMETERS, number_streams)
  F(ASTRONOMY_PARAMETERS, number_stream_parameters) F(ASTRONOMY_PARAMETERS, stream_weights)
  F(ASTRONOMY_PARAMETERS, stream_weight_optimize) F(ASTRONOMY_PARAMETERS, stream_parameters)
  F(ASTRONOMY_PARAMETERS, stream_optimize) F(ASTRONOMY_PARAMETERS, convolve)
  F(ASTRONOMY_PARAMETERS, sgr_coordinates) F(ASTRONOMY_PARAMETERS, aux_bg_profile)
  F(ASTRONOMY_PARAMETERS, wedge) F(ASTRONOMY_PARAMETERS, number_integrals) F(ASTRONOMY_PARAMETERS, integral)
  Z(STAR_POINTS) F(STAR_POINTS, number_stars) F(STAR_POINTS, stars)
  Z(INTEGRAL_AREA) F(INTEGRAL_AREA, mu_steps) F(INTEGRAL_AREA, r_steps) F(INTEGRAL_AREA, mu_min)
  F(INTEGRAL_AREA, r_step_size) F(INTEGRAL_AREA, min_calculation) F(INTEGRAL_AREA, current_calculation)
  F(INTEGRAL_AREA, number_streams) F(INTEGRAL_AREA, background_integral) F(INTEGRAL_AREA, stream_integrals)
  Z(EVALUATION_STATE) F(EVALUATION_STATE, integral) F(EVALUATION_STATE, current_integral)
  F(EVALUATION_STATE, number_integrals) F(EVALUATION_STATE, background_integral)
  F(EVALUATION_STATE, stream_integrals) F(EVALUATION_STATE, current_star_point)
  F(EVALUATION_STATE, num_zero) F(EVALUATION_STATE, bad_jacobians) F(EVALUATION_STATE, prob_sum)
  return 0;
}
"""


def _layout(tmp_path, include, flags):
    src = tmp_path / "probe.c"
    exe = tmp_path / "probe"
    src.write_text(PROBE % {"include": include})
    subprocess.run(["gcc", "-std=gnu99", *flags, str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


@pytest.mark.skipif(not REFERENCE.exists(), reason="reference tree not mounted")
def test_struct_layout_matches_reference_headers(tmp_path):
    ours = _layout(tmp_path, '#include "mw_separation_types.h"', [f"-I{REPO / 'include'}"])
    ref = _layout(tmp_path, '#include "evaluation_state.h"', [f"-I{REFERENCE}"])
    assert ours == ref


def test_ctypes_mirror_matches_header(tmp_path):
    text = _layout(tmp_path, '#include "mw_separation_types.h"', [f"-I{REPO / 'include'}"])
    table = dict(line.rsplit(" ", 1) for line in text.strip().splitlines())
    for key, val in table.items():
        if "." in key:
            t, f = key.split(".")
            assert getattr(getattr(refstructs, t), f).offset == int(val), key
        else:
            assert C.sizeof(getattr(refstructs, key)) == int(val), key
