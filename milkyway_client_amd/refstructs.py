"""ctypes mirrors of the reference evaluator's structs, and owners for them.

The layouts restate ``include/mw_separation_types.h``, which in turn is
byte-identical to the reference's ``milkyway_separation/cpu/parameters.h:27-71``,
``evaluation_state.h:30-58`` and ``star_points.h:27-31``.  The ``*Holder``
classes build the C objects from Python data and keep every buffer they point
into alive for as long as the holder lives.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class INTEGRAL(C.Structure):
    _fields_ = [
        ("r_min", C.c_double), ("r_max", C.c_double), ("r_step_size", C.c_double),
        ("nu_min", C.c_double), ("nu_max", C.c_double), ("nu_step_size", C.c_double),
        ("mu_min", C.c_double), ("mu_max", C.c_double), ("mu_step_size", C.c_double),
        ("r_steps", C.c_int), ("nu_steps", C.c_int), ("mu_steps", C.c_int),
        ("min_calculation", C.c_long), ("max_calculation", C.c_long),
    ]


class ASTRONOMY_PARAMETERS(C.Structure):
    _fields_ = [
        ("parameters_version", C.c_double),
        ("number_background_parameters", C.c_int),
        ("background_weight", C.c_double),
        ("background_parameters", _dp),
        ("background_step", _dp),
        ("background_min", _dp),
        ("background_max", _dp),
        ("background_optimize", _ip),
        ("number_streams", C.c_int),
        ("number_stream_parameters", C.c_int),
        ("stream_weights", _dp),
        ("stream_weight_step", _dp),
        ("stream_weight_min", _dp),
        ("stream_weight_max", _dp),
        ("stream_weight_optimize", _ip),
        ("stream_parameters", C.POINTER(_dp)),
        ("stream_step", C.POINTER(_dp)),
        ("stream_min", C.POINTER(_dp)),
        ("stream_max", C.POINTER(_dp)),
        ("stream_optimize", C.POINTER(_ip)),
        ("convolve", C.c_int),
        ("sgr_coordinates", C.c_int),
        ("aux_bg_profile", C.c_int),
        ("wedge", C.c_int),
        ("number_integrals", C.c_int),
        ("integral", C.POINTER(C.POINTER(INTEGRAL))),
    ]


class STAR_POINTS(C.Structure):
    _fields_ = [("number_stars", C.c_int), ("stars", C.POINTER(_dp))]


class INTEGRAL_AREA(C.Structure):
    _fields_ = [
        ("mu_steps", C.c_int), ("nu_steps", C.c_int), ("r_steps", C.c_int),
        ("mu_min", C.c_double), ("nu_min", C.c_double), ("r_min", C.c_double),
        ("mu_max", C.c_double), ("nu_max", C.c_double), ("r_max", C.c_double),
        ("mu_step_size", C.c_double), ("nu_step_size", C.c_double), ("r_step_size", C.c_double),
        ("min_calculation", C.c_long), ("max_calculation", C.c_long),
        ("current_calculation", C.c_long),
        ("number_streams", C.c_int),
        ("background_integral", C.c_double),
        ("stream_integrals", _dp),
    ]


class EVALUATION_STATE(C.Structure):
    _fields_ = [
        ("integral", C.POINTER(C.POINTER(INTEGRAL_AREA))),
        ("current_integral", C.c_int), ("number_streams", C.c_int),
        ("number_integrals", C.c_int),
        ("background_integral", C.c_double),
        ("stream_integrals", _dp),
        ("current_star_point", C.c_int), ("total_stars", C.c_int),
        ("num_zero", C.c_int),
        ("bad_jacobians", C.c_int),
        ("prob_sum", C.c_double),
    ]


EVAL_FUNC = C.CFUNCTYPE(C.c_double, _dp, C.POINTER(ASTRONOMY_PARAMETERS),
                        C.POINTER(EVALUATION_STATE), C.POINTER(STAR_POINTS))


# --------------------------------------------------------------------------
# Python-side description of a parameter file
# --------------------------------------------------------------------------

@dataclass
class Region:
    """One integration region: r in g magnitudes, mu and nu in degrees."""
    r_min: float
    r_max: float
    r_steps: int
    mu_min: float
    mu_max: float
    mu_steps: int
    nu_min: float
    nu_max: float
    nu_steps: int

    @property
    def cells(self) -> int:
        return self.r_steps * self.mu_steps * self.nu_steps


@dataclass
class Stream:
    weight: float
    parameters: List[float]              # mu deg, r kpc, theta rad, phi rad, sigma kpc
    weight_step: float = 0.1
    weight_min: float = -20.0
    weight_max: float = 20.0
    weight_optimize: int = 1
    step: List[float] = field(default_factory=lambda: [0.3] * 5)
    min: List[float] = field(default_factory=lambda: [-360.0, 0.1, -6.2831853072, -6.2831853072, 0.1])
    max: List[float] = field(default_factory=lambda: [360.0, 60.0, 6.2831853072, 6.2831853072, 20.0])
    optimize: List[int] = field(default_factory=lambda: [1, 1, 1, 1, 1])


@dataclass
class AstronomyParameters:
    background_parameters: List[float]   # alpha, q, r0, delta [, a, b, c]
    streams: List[Stream]
    regions: List[Region]
    convolve: int
    wedge: int
    sgr_coordinates: int = 0
    aux_bg_profile: int = 0
    background_weight: float = 0.0
    parameters_version: float = 0.04
    background_step: List[float] = None
    background_min: List[float] = None
    background_max: List[float] = None
    background_optimize: List[int] = None

    def __post_init__(self):
        nb = len(self.background_parameters)
        if self.background_step is None:
            self.background_step = [0.1] * nb
        if self.background_min is None:
            self.background_min = [0.0] * nb
        if self.background_max is None:
            self.background_max = [30.0] * nb
        if self.background_optimize is None:
            self.background_optimize = [0, 1, 1, 0] + [0] * (nb - 4)

    @property
    def number_streams(self) -> int:
        return len(self.streams)

    def optimized_parameter_count(self) -> int:
        """parameters.c:307-330"""
        n = sum(1 for o in self.background_optimize if o)
        for s in self.streams:
            n += 1 if s.weight_optimize else 0
            n += sum(1 for o in s.optimize if o)
        return n

    def parameter_vector(self) -> np.ndarray:
        """Inverse of set_astronomy_parameters (parameters.c:332-363)."""
        v = [p for p, o in zip(self.background_parameters, self.background_optimize) if o]
        for s in self.streams:
            if s.weight_optimize:
                v.append(s.weight)
            v.extend(p for p, o in zip(s.parameters, s.optimize) if o)
        return np.asarray(v, dtype=np.float64)

    def set_parameter_vector(self, vec: Sequence[float]) -> None:
        """parameters.c:332-363"""
        it = iter(vec)
        for i, o in enumerate(self.background_optimize):
            if o:
                self.background_parameters[i] = float(next(it))
        for s in self.streams:
            if s.weight_optimize:
                s.weight = float(next(it))
            for j, o in enumerate(s.optimize):
                if o:
                    s.parameters[j] = float(next(it))


def _darr(values) -> C.Array:
    vals = list(values)
    return (C.c_double * max(1, len(vals)))(*vals)


def _iarr(values) -> C.Array:
    vals = list(values)
    return (C.c_int * max(1, len(vals)))(*vals)


class ParamsHolder:
    """Owns an ASTRONOMY_PARAMETERS built from an AstronomyParameters."""

    def __init__(self, p: AstronomyParameters):
        self.py = p
        self._keep = []
        k = self._keep
        ap = ASTRONOMY_PARAMETERS()
        ap.parameters_version = p.parameters_version
        ap.number_background_parameters = len(p.background_parameters)
        ap.background_weight = p.background_weight
        for name, vals in (("background_parameters", p.background_parameters),
                           ("background_step", p.background_step),
                           ("background_min", p.background_min),
                           ("background_max", p.background_max)):
            a = _darr(vals)
            k.append(a)
            setattr(ap, name, C.cast(a, _dp))
        a = _iarr(p.background_optimize)
        k.append(a)
        ap.background_optimize = C.cast(a, _ip)

        S = len(p.streams)
        ap.number_streams = S
        ap.number_stream_parameters = 5
        for name, attr in (("stream_weights", "weight"), ("stream_weight_step", "weight_step"),
                           ("stream_weight_min", "weight_min"),
                           ("stream_weight_max", "weight_max")):
            a = _darr(getattr(s, attr) for s in p.streams)
            k.append(a)
            setattr(ap, name, C.cast(a, _dp))
        a = _iarr(s.weight_optimize for s in p.streams)
        k.append(a)
        ap.stream_weight_optimize = C.cast(a, _ip)
        for name, attr in (("stream_parameters", "parameters"), ("stream_step", "step"),
                           ("stream_min", "min"), ("stream_max", "max")):
            rows = (_dp * max(1, S))()
            for i, s in enumerate(p.streams):
                r = _darr(getattr(s, attr))
                k.append(r)
                rows[i] = C.cast(r, _dp)
            k.append(rows)
            setattr(ap, name, C.cast(rows, C.POINTER(_dp)))
        rows = (_ip * max(1, S))()
        for i, s in enumerate(p.streams):
            r = _iarr(s.optimize)
            k.append(r)
            rows[i] = C.cast(r, _ip)
        k.append(rows)
        ap.stream_optimize = C.cast(rows, C.POINTER(_ip))

        ap.convolve = p.convolve
        ap.sgr_coordinates = p.sgr_coordinates
        ap.aux_bg_profile = p.aux_bg_profile
        ap.wedge = p.wedge
        ap.number_integrals = len(p.regions)
        ints = (C.POINTER(INTEGRAL) * max(1, len(p.regions)))()
        for i, g in enumerate(p.regions):
            it = INTEGRAL()
            it.r_min, it.r_max, it.r_steps = g.r_min, g.r_max, g.r_steps
            it.mu_min, it.mu_max, it.mu_steps = g.mu_min, g.mu_max, g.mu_steps
            it.nu_min, it.nu_max, it.nu_steps = g.nu_min, g.nu_max, g.nu_steps
            # parameters.c:185-189
            it.r_step_size = (g.r_max - g.r_min) / float(g.r_steps)
            it.mu_step_size = (g.mu_max - g.mu_min) / float(g.mu_steps)
            it.nu_step_size = (g.nu_max - g.nu_min) / float(g.nu_steps)
            it.min_calculation = 0
            it.max_calculation = g.r_steps * g.mu_steps * g.nu_steps
            k.append(it)
            ints[i] = C.pointer(it)
        k.append(ints)
        ap.integral = C.cast(ints, C.POINTER(C.POINTER(INTEGRAL)))
        self.ap = ap

    @property
    def ptr(self):
        return C.byref(self.ap)

    def region(self, i: int) -> INTEGRAL:
        return self.ap.integral[i].contents

    def stream_weights(self) -> List[float]:
        return [self.ap.stream_weights[i] for i in range(self.ap.number_streams)]


class StarsHolder:
    """Owns a STAR_POINTS whose rows point into one (N, 3) float64 array."""

    def __init__(self, stars: np.ndarray):
        self.array = np.ascontiguousarray(stars, dtype=np.float64).reshape(-1, 3)
        n = self.array.shape[0]
        base = self.array.ctypes.data
        rows = (_dp * max(1, n))()
        addrs = np.uint64(base) + np.uint64(24) * np.arange(n, dtype=np.uint64)
        C.memmove(rows, addrs.ctypes.data, 8 * n)
        self._rows = rows
        sp = STAR_POINTS()
        sp.number_stars = n
        sp.stars = C.cast(rows, C.POINTER(_dp))
        self.sp = sp

    @property
    def ptr(self):
        return C.byref(self.sp)

    def __len__(self):
        return self.array.shape[0]


class StateHolder:
    """Owns an EVALUATION_STATE set up as initialize_state does
    (evaluation_state.c:97-127, :153-176)."""

    def __init__(self, params: ParamsHolder, stars: StarsHolder):
        ap = params.ap
        S = ap.number_streams
        self._keep = []
        es = EVALUATION_STATE()
        es.current_integral = 0
        es.background_integral = 0.0
        si = (C.c_double * max(1, S))()
        self._keep.append(si)
        es.stream_integrals = C.cast(si, _dp)
        es.number_streams = S
        es.total_stars = stars.sp.number_stars
        es.number_integrals = ap.number_integrals
        areas = (C.POINTER(INTEGRAL_AREA) * max(1, ap.number_integrals))()
        for i in range(ap.number_integrals):
            g = ap.integral[i].contents
            ia = INTEGRAL_AREA()
            ia.mu_min, ia.mu_max, ia.mu_steps = g.mu_min, g.mu_max, g.mu_steps
            ia.nu_min, ia.nu_max, ia.nu_steps = g.nu_min, g.nu_max, g.nu_steps
            ia.r_min, ia.r_max, ia.r_steps = g.r_min, g.r_max, g.r_steps
            ia.min_calculation = g.min_calculation
            ia.max_calculation = g.max_calculation
            ia.current_calculation = g.min_calculation
            ia.mu_step_size = (ia.mu_max - ia.mu_min) / ia.mu_steps
            ia.nu_step_size = (ia.nu_max - ia.nu_min) / ia.nu_steps
            ia.r_step_size = (ia.r_max - ia.r_min) / ia.r_steps
            ia.number_streams = S
            a = (C.c_double * max(1, S))()
            self._keep.append(a)
            ia.stream_integrals = C.cast(a, _dp)
            self._keep.append(ia)
            areas[i] = C.pointer(ia)
        self._keep.append(areas)
        es.integral = C.cast(areas, C.POINTER(C.POINTER(INTEGRAL_AREA)))
        self.es = es

    @property
    def ptr(self):
        return C.byref(self.es)

    def area(self, i: int) -> INTEGRAL_AREA:
        return self.es.integral[i].contents

    def stream_integrals(self) -> np.ndarray:
        return np.array([self.es.stream_integrals[i] for i in range(self.es.number_streams)])

    def region_integrals(self, i: int):
        a = self.area(i)
        return a.background_integral, np.array(
            [a.stream_integrals[j] for j in range(a.number_streams)])
