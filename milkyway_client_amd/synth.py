"""Seeded synthetic workloads for the separation likelihood.

Geometry: an SDSS-style wedge (wedge 86, survey great-circle coordinates), the
g-magnitude range [16, 22.5] written as the r axis (the reference's r axis is
in magnitudes: r_kpc = 10**((g - 14.2) / 5)), nu spanning [-1.25, 1.25] deg.

Parameters (per vector): q ~ U[0.4, 0.9], r0 ~ U[4, 8] kpc, alpha = delta = 1,
no auxiliary background profile, background weight 0; per stream
mu ~ U[mu_min, mu_max] deg, r ~ U[5, 40] kpc, theta ~ U[0, pi], phi ~ U[0, 2 pi),
sigma ~ U[1, 5] kpc, weight (log epsilon) ~ U[-4, 0].

Stars: mu, nu and g uniform over the wedge's grid ranges, mapped to (l, b)
through the survey great circle and to kpc through r(g).

RNG: numpy PCG64.  Workload k uses seed 1000 + k; its parameter vector j uses
seed 1000 * (1000 + k) + j.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

from .refstructs import AstronomyParameters, Region, Stream

G_MIN, G_MAX = 16.0, 22.5
NU_MIN, NU_MAX = -1.25, 1.25
WEDGE = 86
MU_MIN = 310.0


def r_kpc_from_g(g):
    return 10.0 ** ((np.asarray(g) - 14.2) / 5.0)


def gc_to_lb(wedge: int, mu, nu):
    """Survey great-circle (mu, nu) deg to galactic (l, b) deg (SDSS stripes):
    rotate about the node at RA 95 deg by the stripe inclination, then the
    fixed equatorial-to-galactic rotation."""
    d2r = np.pi / 180.0
    node = 95.0 * d2r
    eta = wedge * 2.5 - 57.5 - (180.0 if wedge > 46 else 0.0)
    inc = (eta + 32.5) * d2r
    mu = np.asarray(mu, dtype=np.float64) * d2r
    nu = np.asarray(nu, dtype=np.float64) * d2r
    x2 = np.cos(mu - node) * np.cos(nu)
    y2 = np.sin(mu - node) * np.cos(nu)
    z2 = np.sin(nu)
    y1 = y2 * np.cos(inc) - z2 * np.sin(inc)
    z1 = y2 * np.sin(inc) + z2 * np.cos(inc)
    ra = np.arctan2(y1, x2) + node
    dec = np.arcsin(np.clip(z1, -1.0, 1.0))
    rmat = np.array([[-0.054875539726, -0.873437108010, -0.483834985808],
                     [0.494109453312, -0.444829589425, 0.746982251810],
                     [-0.867666135858, -0.198076386122, 0.455983795705]])
    v = np.stack([np.cos(ra) * np.cos(dec), np.sin(ra) * np.cos(dec), np.sin(dec)])
    g = np.tensordot(rmat, v, axes=1)
    l = np.mod(np.arctan2(g[1], g[0]), 2 * np.pi) / d2r
    b = np.arctan2(g[2], np.hypot(g[0], g[1])) / d2r
    return l, b


@dataclass
class WedgeShape:
    r_steps: int
    mu_steps: int
    nu_steps: int
    convolve: int
    n_stars: int
    mu_span: float = 100.0
    cut: tuple = None            # (r_steps, mu_steps, nu_steps, mu_lo, mu_hi) or None

    def regions(self) -> List[Region]:
        regs = [Region(G_MIN, G_MAX, self.r_steps, MU_MIN, MU_MIN + self.mu_span, self.mu_steps,
                       NU_MIN, NU_MAX, self.nu_steps)]
        if self.cut:
            rs, ms, ns, lo, hi = self.cut
            regs.append(Region(G_MIN, G_MAX, rs, lo, hi, ms, NU_MIN, NU_MAX, ns))
        return regs


# The workloads used by the tests and the benchmark.
#   reference-size:  70 x 160 x 64 cells, n = 30, one stream, 10k stars
#   full-size:       700 x 1600 x 640 cells plus a 700 x 400 x 160 cut over 25 deg
#                    of mu, n = 120, three streams, 200k stars
REFERENCE_SIZE = WedgeShape(r_steps=70, mu_steps=160, nu_steps=64, convolve=30, n_stars=10_000)
FULL_SIZE = WedgeShape(r_steps=700, mu_steps=1600, nu_steps=640, convolve=120, n_stars=200_000,
                       cut=(700, 400, 160, MU_MIN + 37.5, MU_MIN + 62.5))
FULL_SIZE_STREAMS = 3

# Uneven batch geometries (nu x mu x r): 64x160x70, 160x400x350, 320x800x700, 640x1600x700.
BATCH_SHAPES = [
    WedgeShape(r_steps=70, mu_steps=160, nu_steps=64, convolve=30, n_stars=10_000),
    WedgeShape(r_steps=350, mu_steps=400, nu_steps=160, convolve=60, n_stars=50_000),
    WedgeShape(r_steps=700, mu_steps=800, nu_steps=320, convolve=120, n_stars=100_000),
    WedgeShape(r_steps=700, mu_steps=1600, nu_steps=640, convolve=120, n_stars=200_000),
]


def random_streams(rng: np.random.Generator, n_streams: int, mu_lo: float, mu_hi: float):
    out = []
    for _ in range(n_streams):
        mu = rng.uniform(mu_lo, mu_hi)
        r = rng.uniform(5.0, 40.0)
        theta = rng.uniform(0.0, np.pi)
        phi = rng.uniform(0.0, 2.0 * np.pi)
        sigma = rng.uniform(1.0, 5.0)
        w = rng.uniform(-4.0, 0.0)
        out.append(Stream(weight=w, parameters=[mu, r, theta, phi, sigma]))
    return out


def make_parameters(shape: WedgeShape, n_streams: int, seed: int) -> AstronomyParameters:
    rng = np.random.Generator(np.random.PCG64(seed))
    q = rng.uniform(0.4, 0.9)
    r0 = rng.uniform(4.0, 8.0)
    regs = shape.regions()
    streams = random_streams(rng, n_streams, regs[0].mu_min, regs[0].mu_max)
    return AstronomyParameters(background_parameters=[1.0, q, r0, 1.0], streams=streams,
                               regions=regs, convolve=shape.convolve, wedge=WEDGE)


def make_stars(shape: WedgeShape, seed: int, n_stars: int = None) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    n = shape.n_stars if n_stars is None else n_stars
    g0 = shape.regions()[0]
    mu = rng.uniform(g0.mu_min, g0.mu_max, n)
    nu = rng.uniform(g0.nu_min, g0.nu_max, n)
    g = rng.uniform(g0.r_min, g0.r_max, n)
    l, b = gc_to_lb(WEDGE, mu, nu)
    return np.stack([l, b, r_kpc_from_g(g)], axis=1)


def workload(k: int, shape: WedgeShape, n_streams: int, n_stars: int = None):
    """Parameters (seed 1000 + k) and stars (seed 1000 + k + 500) of workload k."""
    p = make_parameters(shape, n_streams, 1000 + k)
    stars = make_stars(shape, 1000 + k + 500, n_stars)
    return p, stars


def parameter_vectors(k: int, p: AstronomyParameters, count: int) -> np.ndarray:
    """Parameter vectors j = 0..count-1 of workload k (seed 1000 * (1000 + k) + j),
    laid out as set_astronomy_parameters expects: q, r0, then per stream
    weight, mu, r, theta, phi, sigma."""
    g0 = p.regions[0]
    out = []
    for j in range(count):
        rng = np.random.Generator(np.random.PCG64(1000 * (1000 + k) + j))
        q = rng.uniform(0.4, 0.9)
        r0 = rng.uniform(4.0, 8.0)
        v = [q, r0]
        for s in random_streams(rng, p.number_streams, g0.mu_min, g0.mu_max):
            v += [s.weight] + list(s.parameters)
        out.append(v)
    return np.asarray(out, dtype=np.float64)
