"""Pins of the oracle (CPU, no GPU).

The reference ships no golden vectors (no tests at all), so every pin is
mathematical: dense products of the expanded matrix, brute force on
hand-built trees, algebraic invariants, closed forms, and library routines
(numpy, torch dtype conversions) for the number formats.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, LOG1D, MATERN_CUBE, MATERN_TLR, random_vector
from tests.helpers import rel


@pytest.fixture(scope="module")
def small_mats():
    return {
        "log": HMatrix(LOG1D.with_(n=1024)),
        "galerkin": HMatrix(LOG1D.with_(n=1024, kernel=N.KERNEL_LOG_GALERKIN, geometry=N.GEOM_GRID1D)),
        "matern": HMatrix(MATERN_CUBE.with_(n=2048)),
        "invr": HMatrix(INVR_SPHERE.with_(n=2048)),
        "tlr": HMatrix(MATERN_TLR.with_(n=2048, ntile=256)),
        "usv": HMatrix(INVR_SPHERE.with_(n=2048, lrformat=N.LR_USV)),
    }


def _block_arrays(specs, seed=0):
    """Hand-built leaves -> (blocks, arena); specs: (row, col, m, n, kind, k)."""
    rng = np.random.default_rng(seed)
    bt = N.block_dtype()
    blocks = np.zeros(len(specs), dtype=bt)
    parts, pos = [], 0
    for i, (r, c, m, n, kind, k) in enumerate(specs):
        b = blocks[i]
        b["row_ofs"], b["col_ofs"], b["nrows"], b["ncols"], b["kind"], b["rank"] = r, c, m, n, kind, k
        b["off_s"] = -1
        if kind == N.BLOCK_DENSE:
            a = rng.standard_normal(m * n)
            b["off_a"] = pos; parts.append(a); pos += a.size
        else:
            u, v = rng.standard_normal(m * k), rng.standard_normal(n * k)
            b["off_a"] = pos; parts.append(u); pos += u.size
            b["off_b"] = pos; parts.append(v); pos += v.size
            if kind == N.BLOCK_LOWRANK_SV:
                s = np.sort(rng.uniform(0.1, 2.0, k))[::-1].copy()
                b["off_s"] = pos; parts.append(s); pos += s.size
    return blocks, np.concatenate(parts) if parts else np.zeros(0)


def _brute(specs, blocks, arena, nr, nc):
    A = np.zeros((nr, nc))
    for b in blocks:
        r, c, m, n, k = int(b["row_ofs"]), int(b["col_ofs"]), int(b["nrows"]), int(b["ncols"]), int(b["rank"])
        if int(b["kind"]) == N.BLOCK_DENSE:
            A[r:r + m, c:c + n] += arena[b["off_a"]:b["off_a"] + m * n].reshape(n, m).T
        else:
            U = arena[b["off_a"]:b["off_a"] + m * k].reshape(k, m).T
            V = arena[b["off_b"]:b["off_b"] + n * k].reshape(k, n).T
            if int(b["kind"]) == N.BLOCK_LOWRANK_SV:
                U = U * arena[b["off_s"]:b["off_s"] + k]
            A[r:r + m, c:c + n] += U @ V.T
    return A


# two-level hand-built tree on 10 x 12 (ragged, all three leaf kinds, rank 0)
TINY = [(0, 0, 5, 6, N.BLOCK_DENSE, 0), (0, 6, 2, 6, N.BLOCK_LOWRANK, 2), (2, 6, 3, 6, N.BLOCK_LOWRANK_SV, 3),
        (5, 0, 5, 3, N.BLOCK_LOWRANK, 0), (5, 3, 5, 3, N.BLOCK_DENSE, 0), (5, 6, 5, 6, N.BLOCK_LOWRANK, 4)]


@pytest.mark.parametrize("op", [N.OP_NORMAL, N.OP_TRANSPOSED, N.OP_ADJOINT, N.OP_CONJUGATE])
def test_brute_force_tiny(op):
    blocks, arena = _block_arrays(TINY)
    A = _brute(TINY, blocks, arena, 10, 12)
    trans = op in (N.OP_TRANSPOSED, N.OP_ADJOINT)
    M = A.T if trans else A
    x = np.arange(1.0, M.shape[1] + 1)
    y0 = np.linspace(-1, 1, M.shape[0])
    y = y0.copy()
    oracle.mul_vec(blocks, arena, -0.75, op, x, y)
    assert rel(y, y0 - 0.75 * M @ x) <= 1e-14
    # expansion equals the brute-force sum
    assert rel(oracle.dense_expand(blocks, arena, 10, 12), A) <= 1e-15


@pytest.mark.parametrize("name", ["log", "galerkin", "matern", "invr", "usv"])
def test_mul_vec_vs_dense(small_mats, name):
    hm = small_mats[name]
    A = oracle.dense_expand(hm.blocks, hm.arena, hm.n, hm.n)
    x = random_vector(hm.n, 2)
    for op, M in ((N.OP_NORMAL, A), (N.OP_TRANSPOSED, A.T)):
        y = np.zeros(hm.n)
        oracle.mul_vec(hm.blocks, hm.arena, 1.0, op, x, y)
        assert rel(y, M @ x) <= 1e-14
    # multi-RHS equals column-by-column
    X = np.stack([random_vector(hm.n, s) for s in (3, 4, 5)])
    Y = np.zeros_like(X)
    oracle.mul_vec(hm.blocks, hm.arena, 2.0, N.OP_NORMAL, X, Y)
    assert rel(Y, 2.0 * (A @ X.T).T) <= 1e-14


@pytest.mark.parametrize("name", ["log", "matern", "invr"])
def test_expansion_approximates_kernel(small_mats, name):
    """The generated H-matrix approximates the kernel matrix to about eps."""
    hm = small_mats[name]
    A = oracle.dense_expand(hm.blocks, hm.arena, hm.n, hm.n)
    idx = np.arange(hm.n)
    K = hm.eval(idx, idx)
    assert rel(A, K) <= 10 * hm.problem.eps


def test_invariants(small_mats):
    hm = small_mats["invr"]
    x = random_vector(hm.n, 7)
    w = random_vector(hm.n, 8)
    ax = oracle.mul_vec(hm.blocks, hm.arena, 1.0, N.OP_NORMAL, x, np.zeros(hm.n))
    atw = oracle.mul_vec(hm.blocks, hm.arena, 1.0, N.OP_TRANSPOSED, w, np.zeros(hm.n))
    assert abs(ax @ w - x @ atw) <= 1e-13 * abs(ax @ w)
    # linearity in x and alpha
    a2 = oracle.mul_vec(hm.blocks, hm.arena, 1.0, N.OP_NORMAL, 3 * x - 2 * w, np.zeros(hm.n))
    aw = oracle.mul_vec(hm.blocks, hm.arena, 1.0, N.OP_NORMAL, w, np.zeros(hm.n))
    assert rel(a2, 3 * ax - 2 * aw) <= 1e-14
    # alpha = 0 leaves y untouched (no NaN propagation either)
    y0 = random_vector(hm.n, 9)
    y = y0.copy()
    oracle.mul_vec(hm.blocks, hm.arena, 0.0, N.OP_NORMAL, np.full(hm.n, np.nan), y)
    assert np.array_equal(y, y0)


def test_chunk_and_tlr_orders(small_mats):
    hm = small_mats["invr"]
    x = random_vector(hm.n, 10)
    ys = oracle.mul_vec(hm.blocks, hm.arena, 2.0, N.OP_NORMAL, x, np.zeros(hm.n))
    for nt in (1, 4):
        yc = oracle.mul_vec_chunk(hm.blocks, hm.arena, 2.0, N.OP_NORMAL, x, np.zeros(hm.n), nthreads=nt)
        assert rel(yc, ys) <= 1e-14
    t = small_mats["tlr"]
    x = random_vector(t.n, 11)
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        ys = oracle.mul_vec(t.blocks, t.arena, 1.0, op, x, np.zeros(t.n))
        yt = oracle.tlr_mul_vec(t.blocks, t.arena, 1.0, op, x, np.zeros(t.n), nthreads=3)
        assert rel(yt, ys) <= 1e-14


def test_flops_datasize_hand_counted():
    blocks, _ = _block_arrays(TINY)
    # dense 5x6: 2*30+10 = 70; lr 2x6 k2: 2*2*8+2*8 = 48; lrsv 3x6 k3: 2*3*9+18 = 72;
    # lr 5x3 k0: 0+16 = 16; dense 5x3: 30+10 = 40; lr 5x6 k4: 2*4*11+22 = 110
    assert oracle.mul_vec_flops(blocks) == 70 + 48 + 72 + 16 + 40 + 110
    # bytes: 8*(30 + 2*8 + (3 + 3*9) + 0 + 15 + 4*11) + 16*(10 + 12); the U S V^T
    # leaf counts its k singular values too (matrix/lrsvmatrix.hh:364-370)
    assert oracle.mul_vec_datasize(blocks, 10, 12) == 8 * (30 + 16 + 30 + 0 + 15 + 44) + 16 * 22


# --- precision chooser -------------------------------------------------------

@pytest.mark.parametrize("sv", [[1.0], [3.0, 1e-2, 1e-4, 1e-6], list(np.logspace(0, -7, 20)), [5.0, 4.9, 4.8]])
@pytest.mark.parametrize("eps", [1e-2, 1e-4, 1e-6, 1e-8])
def test_valr_tolerance_root(sv, eps):
    """tol solves sum(1/s) tol^2 - tol + delta = 0 (smaller root), delta =
    eps * s_0; when no real root exists tol = 1 / (2 sum(1/s))."""
    sv = np.array(sv)
    out = oracle.valr_tolerances(sv, eps)
    tol = out * sv
    assert np.allclose(tol, tol[0], rtol=1e-14)
    tol = tol[0]
    S = np.sum(1 / sv)
    delta = eps * sv[0]
    if 4 * delta * S <= 1:
        # smaller root by Vieta (q / larger root): no cancellation; the
        # p/2 - sqrt(.) form loses ~u * p/2 absolutely
        p, q = 1 / S, delta / S
        x = q / (p / 2 + math.sqrt(p * p / 4 - q))
        assert abs(tol - x) <= 1e-14 * p + 1e-14 * x
        assert tol <= 1 / (2 * S) * (1 + 1e-12)        # smaller root
        assert tol >= delta * (1 - 1e-9)                 # tol = delta + S tol^2
    else:
        assert tol == pytest.approx(1 / (2 * S), rel=1e-14)


def test_eps_to_rate_and_formats():
    for m in range(1, 40):
        assert oracle.eps_to_rate(2.0 ** -m) == m
        assert oracle.eps_to_rate(2.0 ** -m * 1.5) == m      # 2^-m < 1.5 2^-m < 2^-(m-1)
    assert oracle.eps_to_rate(0.7) == 1 and oracle.eps_to_rate(3.0) == 1
    assert [oracle.format_for_rate(r) for r in (1, 3, 4, 7, 8, 23, 24, 52)] == [3, 3, 2, 2, 1, 1, 0, 0]


def test_rounding_vs_torch():
    import torch
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.uniform(-4, 2.6, 4000),
                           [0.0, -0.0, 448.0, -448.0, 2 ** -6, 2 ** -9, 2 ** -10, 3 * 2 ** -11, 1.0625, 1.1875]])
    vals = vals[np.abs(vals) < 464]
    t32 = torch.tensor(vals, dtype=torch.float32)
    ref8 = t32.to(torch.float8_e4m3fn).to(torch.float64).numpy()
    got8 = np.array([oracle.round_e4m3(float(v)) for v in t32.numpy()])
    assert np.array_equal(got8, ref8)
    ref16 = t32.to(torch.bfloat16).to(torch.float64).numpy()
    got16 = np.array([oracle.round_bf16(float(v)) for v in t32.numpy()])
    assert np.array_equal(got16, ref16)
    # OCP E4M3 code table: decode of all codes equals torch's
    codes = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float64).numpy()
    for c in range(256):
        d = oracle.decode_e4m3(c)
        assert (math.isnan(d) and math.isnan(codes[c])) or d == codes[c]
    # saturation
    assert oracle.round_e4m3(463.9) == 448.0 and oracle.round_e4m3(1e6) == 448.0


# ---------------------------------------------------------------------------
# block floating point
# ---------------------------------------------------------------------------

BFP_BITS = (8, 16, 25)


@pytest.mark.parametrize("nb", BFP_BITS)
def test_bfp_step_is_minimal(nb):
    """The step 2^k is the smallest power of two with max|v| <= qmax 2^k:
    brute force over k, on random magnitudes and on exact powers of two."""
    rng = np.random.default_rng(nb)
    qmax = 2.0 ** (nb - 1) - 1
    mags = list(np.exp(rng.uniform(-60, 60, 200))) + [2.0 ** e for e in range(-40, 40, 7)] + [qmax, qmax * 2.0 ** -30]
    for a in mags:
        k = oracle.bfp_exp(a, nb)
        assert a <= qmax * 2.0 ** k
        assert a > qmax * 2.0 ** (k - 1)
    assert oracle.bfp_exp(0.0, nb) == 0
    # the step stays a normal FP32 number
    assert oracle.bfp_exp(1e-300, nb) == -126


@pytest.mark.parametrize("nb", BFP_BITS)
def test_bfp_round_is_numpy_rint(nb):
    """Rounding to the nearest multiple of the step, ties to even, is numpy's
    rint (IEEE round-half-even) of v / 2^k, clamped to the NB-bit range."""
    rng = np.random.default_rng(100 + nb)
    qmax = 2.0 ** (nb - 1) - 1
    for _ in range(300):
        k = int(rng.integers(-40, 10))
        v = rng.uniform(-qmax, qmax) * 2.0 ** k
        assert oracle.bfp_round(v, k, nb) == np.rint(v / 2.0 ** k) * 2.0 ** k
    # ties go to the even neighbour, beyond the range to +-qmax
    assert oracle.bfp_round(2.5, 0, nb) == 2.0 and oracle.bfp_round(3.5, 0, nb) == 4.0
    assert oracle.bfp_round(-2.5, 0, nb) == -2.0
    assert oracle.bfp_round(10 * qmax, 0, nb) == qmax and oracle.bfp_round(-10 * qmax, 0, nb) == -qmax


def _bfp_err(col, nb, piece=64):
    return np.linalg.norm(oracle.bfp_round_column(col, nb, piece) - col)


def test_bfp_format_is_the_narrowest_passing():
    """Brute force: the chosen width is the first of 8, 16, 25 bits whose stored
    columns (all of them) meet ||err|| <= tol ||col||, with one step per 64-row
    panel column; columns no width meets go to FP32 / FP64 by rate."""
    rng = np.random.default_rng(7)
    for trial in range(60):
        tol = float(np.exp2(-rng.uniform(1, 30)))
        rate = max(1, int(np.ceil(-np.log2(tol))))
        cols = [rng.standard_normal(int(rng.integers(1, 200))) * np.exp(rng.uniform(-5, 5, 1))
                for _ in range(int(rng.integers(1, 3)))]
        if trial % 3 == 0:   # a spread of magnitudes inside the column
            cols[0] = cols[0] * np.exp(rng.uniform(-8, 0, len(cols[0])))
        f = oracle.bfp_format(cols, tol, rate)
        ok = [all(_bfp_err(c, nb) <= tol * np.linalg.norm(c) for c in cols) for nb in BFP_BITS]
        if any(ok):
            assert f == N.FMT_B8 + ok.index(True), (trial, tol, ok, f)
        else:
            assert f == (N.FMT_F32 if rate <= 23 else N.FMT_F64)


def test_bfp_steps_are_per_panel_column():
    """One step per panel column, not per column: a 64-row panel 10^6 times
    smaller than its neighbour keeps its own relative accuracy, where a
    single step for the whole column would flush it to zero."""
    rng = np.random.default_rng(3)
    col = np.concatenate([rng.standard_normal(64), 1e-6 * rng.standard_normal(64)])
    got = oracle.bfp_round_column(col, 16)
    assert np.linalg.norm(got[64:] - col[64:]) <= 2.0 ** -10 * np.linalg.norm(col[64:])
    assert np.all(oracle.bfp_round_column(col, 16, piece=[128])[64:] == 0.0)
    # exact small integers are stored exactly by the narrowest width
    ints = np.arange(-50.0, 50.0)
    assert oracle.bfp_format([ints], 2.0 ** -60, 60) == N.FMT_B8
    assert np.array_equal(oracle.bfp_round_column(ints, 8), ints)
    # an all-zero column needs nothing
    assert oracle.bfp_format([np.zeros(70)], 2.0 ** -40, 40) == N.FMT_B8
