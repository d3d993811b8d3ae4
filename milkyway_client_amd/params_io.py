"""Readers and writers for the reference's text formats.

* astronomy parameter files: ``milkyway_separation/cpu/parameters.c:107-228``
  (reader) and ``:230-305`` (writer), with the array syntax of
  ``util/io_util.c:30-71,125-141`` (``name[N]: a, b, c``);
* star files: ``milkyway_separation/cpu/star_points.c:69-93`` (a count line,
  then ``l b r`` per line: degrees, degrees, kpc).
"""
from __future__ import annotations

import re
from pathlib import Path
from typing import List, Tuple, Union

import numpy as np

from .refstructs import AstronomyParameters, Region, Stream

_num = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


class ParameterFileError(ValueError):
    pass


class _Lines:
    def __init__(self, text: str):
        self.lines = [ln.strip() for ln in text.splitlines() if ln.strip()]
        self.i = 0

    def peek(self) -> str:
        return self.lines[self.i] if self.i < len(self.lines) else ""

    def take(self, key: str) -> str:
        ln = self.peek()
        if not ln.startswith(key):
            raise ParameterFileError(f"expected '{key}' at line {self.i + 1}, found '{ln}'")
        self.i += 1
        return ln[len(key):]


def _scalar(lines: _Lines, key: str) -> str:
    rest = lines.take(key)
    rest = rest.lstrip()
    if not rest.startswith(":"):
        raise ParameterFileError(f"malformed '{key}' line")
    return rest[1:].strip()


def _array(lines: _Lines, name: str, conv=float) -> List:
    """io_util.c:55-71 / :125-141: name[N]: v0, v1, ..."""
    rest = lines.take(name)
    m = re.match(r"\s*\[(\d+)\]\s*:\s*(.*)$", rest)
    if not m:
        raise ParameterFileError(f"malformed array '{name}'")
    n = int(m.group(1))
    vals = [v for v in re.split(r"[,\s]+", m.group(2).strip()) if v]
    if len(vals) < n:
        raise ParameterFileError(f"Error reading into {name}")
    return [conv(v) for v in vals[:n]]


def _triple(lines: _Lines, key: str) -> Tuple[float, float, int]:
    rest = lines.take(key)
    m = re.match(r"(?:\s*\[\s*\d+\s*\])?\s*:\s*(%s)\s*,\s*(%s)\s*,\s*(\d+)" % (_num, _num), rest)
    if not m:
        raise ParameterFileError(f"malformed '{key}' line")
    return float(m.group(1)), float(m.group(2)), int(m.group(3))


def parse_astronomy_parameters(text: str) -> AstronomyParameters:
    """Parse an astronomy parameter file (parameters.c:107-228)."""
    try:
        return _parse(text)
    except ParameterFileError:
        raise
    except (ValueError, IndexError) as e:
        raise ParameterFileError(str(e)) from e


def _parse(text: str) -> AstronomyParameters:
    L = _Lines(text)
    version = 0.01
    if L.peek().startswith("parameters_version"):
        version = float(_scalar(L, "parameters_version"))
    nbg = int(_scalar(L, "number_parameters"))
    bg_weight = float(_scalar(L, "background_weight"))
    bg = _array(L, "background_parameters")
    bg_step = _array(L, "background_step")
    bg_min = _array(L, "background_min")
    bg_max = _array(L, "background_max")
    bg_opt = _array(L, "optimize_parameter", int)
    # The reference sizes the background arrays from their own [N] prefixes
    # and uses number_parameters as the count everywhere else.
    if len(bg) < nbg:
        raise ParameterFileError("background_parameters shorter than number_parameters")
    ns_txt = _scalar(L, "number_streams")
    m = re.match(r"(\d+)\s*,\s*(\d+)", ns_txt)
    if not m:
        raise ParameterFileError("malformed number_streams")
    n_streams, n_sp = int(m.group(1)), int(m.group(2))
    if n_sp != 5:
        raise ParameterFileError("streams must have 5 parameters")
    streams = []
    for _ in range(n_streams):
        w = float(_scalar(L, "stream_weight"))
        ws = float(_scalar(L, "stream_weight_step"))
        wmin = float(_scalar(L, "stream_weight_min"))
        wmax = float(_scalar(L, "stream_weight_max"))
        wopt = int(_scalar(L, "optimize_weight"))
        sp = _array(L, "stream_parameters")
        st = _array(L, "stream_step")
        smin = _array(L, "stream_min")
        smax = _array(L, "stream_max")
        sopt = _array(L, "optimize_parameter", int)
        streams.append(Stream(weight=w, parameters=sp, weight_step=ws, weight_min=wmin,
                              weight_max=wmax, weight_optimize=wopt, step=st, min=smin,
                              max=smax, optimize=sopt))
    convolve = int(_scalar(L, "convolve"))
    sgr = int(_scalar(L, "sgr_coordinates"))
    aux = 0
    if version > 0.01:
        aux = int(_scalar(L, "aux_bg_profile"))
    wedge = int(_scalar(L, "wedge"))
    r = _triple(L, "r[min,max,steps]")
    mu = _triple(L, "mu[min,max,steps]")
    nu = _triple(L, "nu[min,max,steps]")
    regions = [Region(r[0], r[1], r[2], mu[0], mu[1], mu[2], nu[0], nu[1], nu[2])]
    n_cuts = int(_scalar(L, "number_cuts"))
    for _ in range(n_cuts):
        rc = _triple(L, "r_cut[min,max,steps]")
        mc = _triple(L, "mu_cut[min,max,steps]")
        nc = _triple(L, "nu_cut[min,max,steps]")
        regions.append(Region(rc[0], rc[1], rc[2], mc[0], mc[1], mc[2], nc[0], nc[1], nc[2]))
    return AstronomyParameters(
        background_parameters=bg[:nbg], streams=streams, regions=regions,
        convolve=convolve, wedge=wedge, sgr_coordinates=sgr, aux_bg_profile=aux,
        background_weight=bg_weight, parameters_version=version,
        background_step=bg_step[:nbg], background_min=bg_min[:nbg],
        background_max=bg_max[:nbg], background_optimize=bg_opt[:nbg])


def read_astronomy_parameters(path: Union[str, Path]) -> AstronomyParameters:
    return parse_astronomy_parameters(Path(path).read_text())


def _fmt_darr(name: str, vals) -> str:
    """io_util.c:30-39"""
    return "%s[%d]: %s" % (name, len(vals), ", ".join("%.20f" % v for v in vals))


def _fmt_iarr(name: str, vals) -> str:
    """io_util.c:41-50"""
    return "%s[%d]:  %s" % (name, len(vals), ", ".join("%d" % v for v in vals))


def format_astronomy_parameters(p: AstronomyParameters) -> str:
    """parameters.c:230-305"""
    out = ["parameters_version: %f" % p.parameters_version,
           "number_parameters: %d" % len(p.background_parameters),
           "background_weight: %f" % p.background_weight,
           _fmt_darr("background_parameters", p.background_parameters),
           _fmt_darr("background_step", p.background_step),
           _fmt_darr("background_min", p.background_min),
           _fmt_darr("background_max", p.background_max),
           _fmt_iarr("optimize_parameter", p.background_optimize),
           "number_streams: %d, %d" % (len(p.streams), 5)]
    for s in p.streams:
        out += ["stream_weight: %f" % s.weight,
                "stream_weight_step: %f" % s.weight_step,
                "stream_weight_min: %f" % s.weight_min,
                "stream_weight_max: %f" % s.weight_max,
                "optimize_weight: %d" % s.weight_optimize,
                _fmt_darr("stream_parameters", s.parameters),
                _fmt_darr("stream_step", s.step),
                _fmt_darr("stream_min", s.min),
                _fmt_darr("stream_max", s.max),
                _fmt_iarr("optimize_parameter", s.optimize)]
    out += ["convolve: %d" % p.convolve,
            "sgr_coordinates: %d" % p.sgr_coordinates,
            "aux_bg_profile: %d" % p.aux_bg_profile,
            "wedge: %d" % p.wedge]
    g = p.regions[0]
    out += ["r[min,max,steps]: %f, %f, %d" % (g.r_min, g.r_max, g.r_steps),
            "mu[min,max,steps]: %f, %f, %d" % (g.mu_min, g.mu_max, g.mu_steps),
            "nu[min,max,steps]: %f, %f, %d" % (g.nu_min, g.nu_max, g.nu_steps),
            "number_cuts: %d" % (len(p.regions) - 1)]
    for g in p.regions[1:]:
        out += ["r_cut[min,max,steps][3]: %f, %f, %d" % (g.r_min, g.r_max, g.r_steps),
                "mu_cut[min,max,steps][3]: %f, %f, %d" % (g.mu_min, g.mu_max, g.mu_steps),
                "nu_cut[min,max,steps][3]: %f, %f, %d" % (g.nu_min, g.nu_max, g.nu_steps)]
    return "\n".join(out) + "\n"


def write_astronomy_parameters(path: Union[str, Path], p: AstronomyParameters) -> None:
    Path(path).write_text(format_astronomy_parameters(p))


def parse_star_points(text: str) -> np.ndarray:
    """star_points.c:69-81: count line, then 'l b r' rows."""
    toks = text.split()
    if not toks:
        raise ParameterFileError("empty star file")
    n = int(toks[0])
    vals = np.asarray(toks[1:1 + 3 * n], dtype=np.float64)
    if vals.size != 3 * n:
        raise ParameterFileError(f"star file declares {n} stars, holds {vals.size // 3}")
    return vals.reshape(n, 3)


def read_star_points(path: Union[str, Path], limit: int = None) -> np.ndarray:
    with open(path) as f:
        n = int(f.readline().split()[0])
        if limit is not None:
            n = min(n, limit)
        rows = np.loadtxt(f, dtype=np.float64, max_rows=n, ndmin=2)
    if rows.shape[0] != n:
        raise ParameterFileError(f"star file declares {n} stars, holds {rows.shape[0]}")
    return rows[:, :3]


def format_star_points(stars: np.ndarray) -> str:
    """star_points.c:83-93 (%lf, six decimals)."""
    lines = ["%d" % stars.shape[0]]
    lines += ["%f %f %f" % (l, b, r) for l, b, r in stars]
    return "\n".join(lines) + "\n"


def write_star_points(path: Union[str, Path], stars: np.ndarray) -> None:
    Path(path).write_text(format_star_points(stars))
