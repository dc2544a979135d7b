"""Exact-arithmetic fixtures on the GPU: the plans in every storage format
(FP64; FP32; every low-rank column forced to BF16, to FP8-E4M3 with its
power-of-two scale, or to 8-, 16- or 25-bit block floating point) must return the fixture's integer results bit for bit,
for the normal and transposed products, three right-hand sides on the
matrix cores, FP32 vectors, the one-level TLR entry point and a batch.
Every partial sum on these inputs is an integer below 2^24 (checked on the
CPU in test_golden.py), so FP32 accumulation in any order is exact and
any difference is a defect, not rounding."""
from __future__ import annotations

import numpy as np
import pytest

from hlr_amd import _native as N
from hlr_amd.batch import Batch
from hlr_amd.plan import Plan
from tests.golden_data import CASES, load

pytestmark = pytest.mark.gpu

CONFIGS = [("fp64", None), ("fp32", None), ("valr", "bf16"), ("valr", "fp8_e4m3"), ("bfp", "bfp8"), ("bfp", "bfp16"),
           ("bfp", "bfp25")]


def _plan(g, prec, force):
    return Plan(g["blocks"], g["arena"], g["nrows"], g["ncols"], precision=prec, eps=1e-6, force_format=force,
                dense_prec="fp64" if prec == "fp64" else "fp32")


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("prec,force", CONFIGS)
def test_exact(cuda, case, prec, force):
    import torch
    g = load(case)
    print(f"[gpu] golden {case} {prec} {force}", flush=True)
    p = _plan(g, prec, force)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=cuda)
    y = t(g["y0"])
    p.apply(g["alpha"], N.OP_NORMAL, t(g["x"]), y)
    assert np.array_equal(y.cpu().numpy(), g["y"])
    yt = t(g["yt0"])
    p.apply(g["alpha"], N.OP_TRANSPOSED, t(g["xt"]), yt)
    assert np.array_equal(yt.cpu().numpy(), g["yt"])
    Y = t(g["Y0"])
    p.apply(g["alpha"], N.OP_NORMAL, t(g["X"]), Y)
    assert np.array_equal(Y.cpu().numpy(), g["Y"])
    # 32 right-hand sides (an even count: staged in pairs), the golden
    # columns repeated in a mixed order; every product is exact, so each
    # column must equal its golden result bit for bit
    idx = np.arange(32) * 5 % len(g["X"])
    Y = t(g["Y0"][idx])
    p.apply(g["alpha"], N.OP_NORMAL, t(g["X"][idx]), Y)
    assert np.array_equal(Y.cpu().numpy(), g["Y"][idx])
    y32 = t(g["y0"], torch.float32)
    p.apply(g["alpha"], N.OP_NORMAL, t(g["x"], torch.float32), y32)
    assert np.array_equal(y32.cpu().numpy().astype(np.float64), g["y"])
    if case == "tlr":
        y = t(g["y0"])
        p.tlr_apply(g["alpha"], N.OP_NORMAL, t(g["x"]), y)
        assert np.array_equal(y.cpu().numpy(), g["y"])
        yt = t(g["yt0"])
        p.tlr_apply(g["alpha"], N.OP_TRANSPOSED, t(g["xt"]), yt)
        assert np.array_equal(yt.cpu().numpy(), g["yt"])


@pytest.mark.parametrize("prec,force", [("fp64", None), ("valr", "fp8_e4m3"), ("bfp", "bfp8")])
def test_exact_batch(cuda, prec, force):
    import torch
    gs = [load(c) for c in CASES]
    print(f"[gpu] golden batch {prec} {force}", flush=True)
    plans = [_plan(g, prec, force) for g in gs]
    # member 0 with one right-hand side, member 1 with three
    xs = [torch.as_tensor(gs[0]["x"], device=cuda), torch.as_tensor(gs[1]["X"], device=cuda)]
    ys = [torch.as_tensor(gs[0]["y0"], device=cuda), torch.as_tensor(gs[1]["Y0"], device=cuda)]
    b = Batch(plans, [1, 3])
    b.apply(xs, ys, [gs[0]["alpha"], gs[1]["alpha"]])
    b.check()
    assert np.array_equal(ys[0].cpu().numpy(), gs[0]["y"])
    assert np.array_equal(ys[1].cpu().numpy(), gs[1]["Y"])
    b.close()
