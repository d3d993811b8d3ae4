"""Frozen algorithmic work model of one likelihood evaluation.

Counted per convolution point from the integrand's source expressions
(evaluation_optimized.c:222-296): every add, subtract and multiply is one
flop; each transcendental is charged a fixed flop-equivalent weight.

* background: 15 ops + 1 sqrt + 1 divide          -> 15 + 10 + 10 = 35
* each stream: 22 ops + 1 exp                      -> 22 + 20      = 42
* star-pass points also rebuild their tables (pow, exp, ~13 ops)  -> +53

These weights are fixed once and never retuned against measurements.
Byte counts are the algorithmic HBM traffic of one launch: the column
geometry (4 doubles per column), the r-step tables (2 doubles per
(r-step, convolution point) plus one 64-byte row per r-step) and, for the
star pass, 5 doubles per star read plus one double written per star.
"""
from __future__ import annotations

W_EXP = 20
W_SQRT = 10
W_DIV = 10
BG_OPS = 15
STREAM_OPS = 22
STAR_TABLE_EXTRA = 53


def flops_per_point(n_streams: int) -> int:
    return BG_OPS + W_SQRT + W_DIV + n_streams * (STREAM_OPS + W_EXP)


def integral_points(regions, n_conv: int) -> int:
    """Convolution points of the integral over every region (r x mu x nu x n)."""
    return sum(g.r_steps * g.mu_steps * g.nu_steps for g in regions) * n_conv


def star_points(n_stars: int, n_conv: int) -> int:
    return n_stars * n_conv


def integral_flops(regions, n_conv: int, n_streams: int) -> float:
    return float(integral_points(regions, n_conv)) * flops_per_point(n_streams)


def star_flops(n_stars: int, n_conv: int, n_streams: int) -> float:
    return float(star_points(n_stars, n_conv)) * (flops_per_point(n_streams) + STAR_TABLE_EXTRA)


def evaluation_flops(regions, n_conv: int, n_streams: int, n_stars: int) -> float:
    return integral_flops(regions, n_conv, n_streams) + star_flops(n_stars, n_conv, n_streams)


def integral_bytes(regions, n_conv: int) -> int:
    b = 0
    for g in regions:
        b += g.mu_steps * g.nu_steps * 4 * 8           # column geometry
        b += g.r_steps * n_conv * 2 * 8 + g.r_steps * 64   # r-step tables
    return b


def star_bytes(n_stars: int) -> int:
    return n_stars * (5 * 8 + 8)
