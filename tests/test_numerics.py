"""Arithmetic identities the device code relies on, checked exhaustively in exact rational arithmetic.

The tracer's attribute decoding divides 8-bit unorm values by 255 and 16-bit snorm values by 32767 (unpack4x8unorm,
unpack2x16snorm). The device forms x / d as q = x * r, q' = fma(fma(-q, d, x), r, q) with r = f32(1 / d) (vmath.h
div255 / div32767). Over every input these operands can take, that must be the correctly rounded quotient, which is
what the oracle's IEEE division gives."""
from fractions import Fraction as F
import struct

import numpy as np


def _bits(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


def rnd(q):
    """q (a Fraction) rounded to the nearest f32, ties to even."""
    if q == 0:
        return F(0)
    c = np.float32(float(q))
    cands = [np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))]
    best = min(cands, key=lambda v: (abs(F(float(v)) - q), _bits(v) & 1))
    return F(float(best))


def fast_quotient_mismatches(d, xs):
    r = rnd(F(1, d))
    bad = []
    for x in xs:
        X = F(x)
        q0 = rnd(X * r)
        q = rnd(rnd(-q0 * d + X) * r + q0)  # each fma rounds once
        if q != rnd(X / d):
            bad.append(x)
    return bad


def test_rounding_helper_against_numpy():
    rng = np.random.default_rng(1)
    for a, b in rng.integers(-10 ** 6, 10 ** 6, (2000, 2)):
        if b == 0:
            continue
        assert rnd(F(int(a), int(b))) == F(float(np.float32(np.float64(a) / np.float64(b))))


def test_unorm8_division_is_exact():
    assert fast_quotient_mismatches(255, range(256)) == []


def test_snorm16_division_is_exact():
    assert fast_quotient_mismatches(32767, range(-32768, 32768)) == []


def test_small_range_cell_decomposition():
    """The voxelisation and brick marking split the cell index q of a small range (at most 8 cells, extents nx, ny
    <= 8) into x, y, z with q / n = (q * m) >> 12, m = 4096 // n + 1, in place of divisions (build.hip
    cell_of_small). Exhaustive over every extent and every q < 64 (a small range has q < 8)."""
    for nx in range(1, 9):
        for ny in range(1, 9):
            mx, my = 4096 // nx + 1, 4096 // ny + 1
            for q in range(64):
                t = (q * mx) >> 12
                z = (t * my) >> 12
                assert (q - t * nx, t - z * ny, z) == (q % nx, (q // nx) % ny, q // (nx * ny))
