"""Builds libmw_separation_hip.so in-tree for gfx950.

The HIP sources are compiled with hipcc; the host-side table builder with g++
and -ffp-contract=off, so that its FP64 arithmetic rounds like the
reference's scalar code.  All HIP code is compiled with packed-FP32
instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
LIB_NAME = "libmw_separation_hip.so"
LIB_PATH = LIB_DIR / LIB_NAME
ARCH = os.environ.get("MW_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["mw_kernels.hip", "mw_api.hip", "mw_batch.hip", "mw_server.hip", "mw_separation.hip",
               "mw_search.hip", "mw_quad.hip", "mw_noalloc.hip"]
HOST_SOURCES = ["mw_tables.cpp"]
C_SOURCES = ["mw_io.c", "mw_search_host.c"]
CLIENT_SOURCE = "client/mw_client.c"
CLIENT_NAME = "milkyway_separation_hip"
CLIENT_PATH = LIB_DIR / CLIENT_NAME
SEPARATE_SOURCE = "client/mw_separate.c"
SEPARATE_NAME = "separation_hip"
ERRORS_SOURCE = "client/mw_errors.c"
ERRORS_NAME = "errors_hip"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
]
GXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             "-pthread"]
GCC_FLAGS = ["-O2", "-std=c99", "-D_GNU_SOURCE", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(map(str, cmd)) + "\n" + r.stdout
                           + r.stderr)
    return r


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library (always rebuilds)."""
    hipcc = _hipcc()
    obj_dir = LIB_DIR / "obj"
    obj_dir.mkdir(parents=True, exist_ok=True)
    inc = [f"-I{REPO / 'include'}", f"-I{CSRC}"]
    jobs = []
    for src in HIP_SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        jobs.append([hipcc, *HIPCC_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        jobs.append(["g++", *GXX_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
    for src in C_SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        jobs.append(["gcc", *GCC_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr, file=sys.stderr)
    objs = [str(obj_dir / (Path(s).stem + ".o")) for s in HIP_SOURCES + HOST_SOURCES + C_SOURCES]
    tmp = LIB_DIR / (LIB_NAME + ".tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB_PATH)
    # Command-line programs, linked against the library with an $ORIGIN rpath
    # so that they move together with it.
    for src, name in ((CLIENT_SOURCE, CLIENT_NAME), (SEPARATE_SOURCE, SEPARATE_NAME),
                      (ERRORS_SOURCE, ERRORS_NAME)):
        tmp = LIB_DIR / (name + ".tmp")
        _run(["gcc", *GCC_FLAGS, *inc, str(CSRC / src), "-o", str(tmp), f"-L{LIB_DIR}",
              "-lmw_separation_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, LIB_DIR / name)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
