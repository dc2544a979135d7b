"""hmvm_plan_arena_source: the plan layout as a map from stored values back
to the leaves' own FP64 arena (CPU, host-only plans).

Pinned without unpack_block: every stored value equals the input value at
its source offset (exactly for FP64, and after the format's rounding for
FP32 and BF16), and the sources of all formats together cover every factor
value of every leaf exactly once (S values of U S V^T leaves stay outside
the arenas).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, LOG1D, MATERN_TLR
from hlr_amd.plan import Plan
from tests.golden_data import load


def _factor_offsets(blocks):
    """offsets of every U, V and D value in the input arena"""
    out = []
    for b in blocks:
        m, n, k = int(b["nrows"]), int(b["ncols"]), int(b["rank"])
        if int(b["kind"]) == N.BLOCK_DENSE:
            out.append(np.arange(m * n) + int(b["off_a"]))
        elif k:
            out.append(np.arange(m * k) + int(b["off_a"]))
            out.append(np.arange(n * k) + int(b["off_b"]))
    return np.sort(np.concatenate(out)) if out else np.zeros(0, dtype=np.int64)


def _cases():
    from types import SimpleNamespace
    g = load("h_ragged")
    ex = SimpleNamespace(blocks=g["blocks"], arena=g["arena"], nrows=g["nrows"], ncols=g["ncols"])
    return [("exact_h", ex), ("log1d", HMatrix(LOG1D)), ("sphere", HMatrix(INVR_SPHERE.with_(n=4096))),
            ("tlr", HMatrix(MATERN_TLR.with_(n=2048)))]


@pytest.fixture(scope="module")
def cases():
    return _cases()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "valr", "bfp"])
def test_arena_source_covers_and_matches(cases, prec):
    for name, hm in cases:
        plan = Plan(hm.blocks, hm.arena, hm.nrows, hm.ncols, precision=prec, eps=1e-6, device=-1)
        seen = []
        for f in range(N.NFMT):
            src = plan.arena_source(f)
            assert src.size == plan.arena_elems(f)
            if src.size == 0:
                continue
            live = src >= 0
            assert src[live].max() < hm.arena.size, name
            seen.append(src[live])
            vals = hm.arena[src[live]]
            if f == N.FMT_F64:
                stored = plan.arena_numpy(f)[:src.size]
                assert np.array_equal(stored[live], vals), (name, prec)
                assert np.all(stored[~live] == 0), (name, prec)
            elif f == N.FMT_F32 and prec == "fp32":
                stored = plan.arena_numpy(f)[:src.size]
                assert np.array_equal(stored[live], vals.astype(np.float32)), (name, prec)
            elif f == N.FMT_BF16:
                stored = plan.arena_numpy(f)[:src.size].astype(np.uint32) << 16
                ref = torch.from_numpy(vals.astype(np.float32)).to(torch.bfloat16).float().numpy()
                assert np.array_equal(stored[live].view(np.float32), ref), (name, prec)
        allsrc = np.sort(np.concatenate(seen))
        assert np.array_equal(allsrc, _factor_offsets(hm.blocks)), (name, prec)


def test_arena_source_rejects_bad_format(cases):
    _, hm = cases[0]
    plan = Plan(hm.blocks, hm.arena, hm.nrows, hm.ncols, precision="fp64", device=-1)
    assert N.hmvm().hmvm_plan_arena_source(plan.handle, N.NFMT, None) == 1
    assert N.hmvm().hmvm_plan_arena_source(plan.handle, -1, None) == 1
    assert N.hmvm().hmvm_plan_arena_source(None, 0, None) == 1
