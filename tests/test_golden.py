"""Exact-arithmetic fixtures (CPU): hand-built ragged H and TLR trees with
small-integer factors and vectors whose products were computed with Python
integers (tests/golden/make_golden.py).  Every partial sum is an integer
below 2^24, so any correct evaluation -- FP64 or FP32 accumulation, any
order -- returns the fixture's numbers exactly; these tests require
equality, not a tolerance."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.plan import Plan
from tests.golden_data import CASES, abs_bound, exact_block, load


@pytest.fixture(scope="module", params=CASES)
def g(request):
    return load(request.param)


def test_sums_stay_exact_in_fp32(g):
    # the precondition that makes FP32 accumulation exact on these inputs
    assert abs_bound(g, g["x"], False) < 2 ** 24
    assert abs_bound(g, g["xt"], True) < 2 ** 24
    for x in g["X"]:
        assert abs_bound(g, x, False) < 2 ** 24


def test_oracle_exact(g):
    a = g["alpha"]
    y = g["y0"].copy()
    oracle.mul_vec(g["blocks"], g["arena"], a, N.OP_NORMAL, g["x"], y)
    assert np.array_equal(y, g["y"])
    yt = g["yt0"].copy()
    oracle.mul_vec(g["blocks"], g["arena"], a, N.OP_TRANSPOSED, g["xt"], yt)
    assert np.array_equal(yt, g["yt"])
    Y = g["Y0"].copy()
    oracle.mul_vec(g["blocks"], g["arena"], a, N.OP_NORMAL, g["X"], Y)
    assert np.array_equal(Y, g["Y"])
    for op, x, y0, exp in ((N.OP_NORMAL, g["x"], g["y0"], g["y"]), (N.OP_TRANSPOSED, g["xt"], g["yt0"], g["yt"])):
        y = y0.copy()
        oracle.mul_vec_chunk(g["blocks"], g["arena"], a, op, x, y, nthreads=4)
        assert np.array_equal(y, exp)
    A = oracle.dense_expand(g["blocks"], g["arena"], g["nrows"], g["ncols"])
    assert np.array_equal(g["y0"] + a * (A @ g["x"]), g["y"])


def test_oracle_tlr_exact():
    g = load("tlr")
    for nt in (1, 4):
        y = g["y0"].copy()
        oracle.tlr_mul_vec(g["blocks"], g["arena"], g["alpha"], N.OP_NORMAL, g["x"], y, nthreads=nt)
        assert np.array_equal(y, g["y"])
        yt = g["yt0"].copy()
        oracle.tlr_mul_vec(g["blocks"], g["arena"], g["alpha"], N.OP_TRANSPOSED, g["xt"], yt, nthreads=nt)
        assert np.array_equal(yt, g["yt"])


@pytest.mark.parametrize("prec,force", [("fp64", None), ("fp32", None), ("valr", "bf16"), ("valr", "fp8_e4m3"),
                                        ("bfp", "bfp8"), ("bfp", "bfp16"), ("bfp", "bfp25")])
def test_layout_holds_integers_exactly(g, prec, force):
    """Small integers are exact in every storage format (FP8-E4M3 after the
    power-of-two column scale), so unpacking the stored plan must give the
    integer blocks back bit for bit.  S of a U diag(S) V^T leaf is not part
    of the stored factors (the plan applies it in FP64 to V^T x)."""
    p = Plan(g["blocks"], g["arena"], g["nrows"], g["ncols"], precision=prec, eps=1e-6, force_format=force,
             dense_prec="fp64" if prec == "fp64" else "fp32", device=-1)
    ar = [p.arena_numpy(f) for f in range(N.NFMT)]
    for b in range(len(g["blocks"])):
        A, B = p.unpack_block(b, ar)
        blk = g["blocks"][b]
        if int(blk["kind"]) == N.BLOCK_LOWRANK_SV:
            A = A * g["arena"][blk["off_s"]:blk["off_s"] + int(blk["rank"])]
        got = A if B is None else A @ B.T
        assert np.array_equal(got, exact_block(g, b)), (prec, force, b)
