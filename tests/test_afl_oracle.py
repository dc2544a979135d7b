"""The oracle's adaptive float (AFL) codec (compress/afl.hh:220-420 with
fixed code widths of 12, 16 and 24 bits, 5 exponent bits), pinned without
restating its bit operations:
  - values on the code grid s (2^E (1 + j 2^-m) - 1) decode exactly;
  - any other value decodes to the largest grid value not above it in
    magnitude (truncation), found by brute force over the grid;
  - the scale is the column's smallest non-zero magnitude while the
    column's range fits 2^32, and max / 2^32 (rounded down to FP32) beyond;
    values below the scale and zeros take the all-ones code and decode to 0;
  - the sign is kept, and 1 + 5 + m bits hold every code.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle

WIDTHS = {12: 6, 16: 10, 24: 18}      # code bits -> mantissa bits


def _grid(s, m, emax=32):
    """every positive value the codes of mantissa m stand for (exact in FP64
    for E <= 23; above, FP32 rounds a - 1 as the decoder does)"""
    out = []
    for E in range(1, emax + 1):
        for j in range(2 ** m):
            a = np.float32(2.0 ** E * (1 + j / 2 ** m))
            out.append(float(np.float32(a - np.float32(1.0))) * s)
    return np.unique(np.array(out))


@pytest.mark.parametrize("nb", [12, 16])
def test_grid_values_exact_and_truncation(nb):
    m = WIDTHS[nb]
    s = 0.375
    grid = _grid(s, m, emax=12)
    rng = np.random.default_rng(nb)
    # grid points decode exactly, with either sign
    for g in rng.choice(grid, 200):
        for v in (g, -g):
            c = oracle.afl_encode(v, s, nb)
            assert 0 <= c < 2 ** nb
            assert oracle.afl_decode(c, s, nb) == v
    # other values: the largest grid value not above |v|
    for v in rng.uniform(s, grid[-1], 300):
        sgn = -1.0 if rng.random() < 0.5 else 1.0
        got = oracle.afl_decode(oracle.afl_encode(sgn * v, s, nb), s, nb)
        want = grid[np.searchsorted(grid, v, side="right") - 1]
        assert got == sgn * want, (v, got, want)


def test_zero_and_small_values_and_widths():
    for nb, m in WIDTHS.items():
        zero = 2 ** nb - 1
        assert oracle.afl_encode(0.0, 1.0, nb) == zero
        assert oracle.afl_decode(zero, 1.0, nb) == 0.0
        assert oracle.afl_encode(0.5, 1.0, nb) == zero          # below the scale
        assert oracle.afl_decode(oracle.afl_encode(1.0, 1.0, nb), 1.0, nb) == 1.0   # a = 2
        top = 2.0 ** 32                                           # a = 2^32 + 1, the largest range
        assert abs(oracle.afl_decode(oracle.afl_encode(top, 1.0, nb), 1.0, nb) - top) <= top * 2.0 ** -m


def test_scale_rule():
    col = np.array([0.0, 3e-3, -2.5, 1e-2, 0.0, -7e-4])
    assert oracle.afl_scale(col) == np.float32(7e-4) or oracle.afl_scale(col) == np.nextafter(np.float32(7e-4), np.float32(0))
    assert oracle.afl_scale(col) <= 7e-4
    wide = np.array([1.0, 1e-12, -3.0])                          # range beyond 2^32
    s = oracle.afl_scale(wide)
    assert s <= 3.0 * 2.0 ** -32 and s > 3.0 * 2.0 ** -32 * (1 - 2 ** -23)
    assert oracle.afl_scale(np.zeros(5)) == 1.0
    # the column's values decode within 2^-m (|v| + s) of themselves
    rng = np.random.default_rng(1)
    for nb, m in WIDTHS.items():
        c = rng.standard_normal(256) * np.exp(rng.uniform(-6, 0, 256))
        r = oracle.afl_round_column(c, nb, 256)
        s = oracle.afl_scale(c)
        assert np.all(np.abs(r - c) <= 2.0 ** -m * (np.abs(c) + s) * (1 + 1e-7))
        assert np.all(np.abs(r) <= np.abs(c))


def test_format_is_the_narrowest_passing():
    """afl_format against brute force over the pinned rounding: the first
    width whose stored pieces keep every column within tol, else FP32 for
    rate <= 23 and FP64 beyond."""
    from hlr_amd import _native as N
    rng = np.random.default_rng(7)
    seen = set()
    for trial in range(60):
        n = int(rng.integers(1, 200))
        cols = [rng.standard_normal(n) * np.exp(rng.uniform(-8, 0, n)) for _ in range(int(rng.integers(1, 3)))]
        rate = int(rng.integers(2, 30))
        tol = 2.0 ** -rate * rng.uniform(1.0, 2.0)
        want = N.FMT_F32 if rate <= 23 else N.FMT_F64
        for f, nb in ((N.FMT_A12, 12), (N.FMT_A16, 16), (N.FMT_A24, 24)):
            ok = all(np.linalg.norm(oracle.afl_round_column(c, nb) - c) <= tol * np.linalg.norm(c) for c in cols)
            if ok:
                want = f
                break
        got = oracle.afl_format(cols, tol, rate)
        assert got == want, (trial, got, want)
        seen.add(got)
    assert {N.FMT_A12, N.FMT_A16, N.FMT_A24} <= seen
    # columns the 6-bit code holds exactly (grid values of one piece) pass at any tol
    g = np.array([1.0, 2.0 * (1 + 1 / 64) - 1, 3.0, 7.0, -15.0])
    assert oracle.afl_format([g], 2.0 ** -50, 50) == N.FMT_A12
    assert oracle.afl_format([np.zeros(70)], 2.0 ** -40, 40) == N.FMT_A12
