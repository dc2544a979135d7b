"""Uniform-H and H^2 products on the GPU (hmvm_uplan_apply through the
C-ABI) against the oracle's restatement of hlr::uniform::mul_vec and
hlr::h2::mul_vec on the same matrix.

Bars: FP64 plans <= 1e-13 relative to the oracle, forward and transposed,
with alpha and accumulation; FP32 storage <= 2e-6; mixed precision (VALR
formats chosen from the bases' singular values) <= eps against the FP64
oracle.  Matrices: the uniform-H and H^2 forms of 1/r on the sphere and
Matern in the cube (n = 16384 and 8192), converted at eps = 1e-6.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, MATERN_CUBE, random_vector
from hlr_amd.uniform import UPlan, build_uniform
from tests.helpers import rel

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


@pytest.fixture(scope="module")
def mats():
    out = {}
    for name, prob in (("sphere", INVR_SPHERE.with_(n=16384)), ("matern", MATERN_CUBE.with_(n=8192))):
        hm = HMatrix(prob)
        out[name] = {"uniform": build_uniform(hm, 1e-6, nested=False), "h2": build_uniform(hm, 1e-6, nested=True)}
    return out


def _gpu(plan, alpha, op, x, y0, cuda):
    import torch
    xt = torch.as_tensor(x, device=cuda)
    yt = torch.as_tensor(y0, device=cuda).clone()
    plan.apply(alpha, op, xt, yt)
    torch.cuda.synchronize()
    return yt.cpu().numpy()


@pytest.mark.parametrize("name", ["sphere", "matern"])
@pytest.mark.parametrize("form", ["uniform", "h2"])
def test_fp64(cuda, mats, name, form):
    um = mats[name][form]
    plan = UPlan(um, precision="fp64")
    info = plan.info()
    print(f"[gpu] uplan {name} {form} n={um.nrows} levels={info['levels']} bytes={info['datasize']:.3g}", flush=True)
    x = random_vector(um.ncols, 3)
    y0 = random_vector(um.nrows, 4)
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        y = _gpu(plan, 2.5, op, x, y0, cuda)
        yr = y0.copy()
        oracle.uniform_mul_vec(um.rowcb, um.colcb, um.blocks, um.arena, 2.5, op, x, yr)
        e = rel(y, yr)
        print(f"[gpu] uplan {name} {form} op={op} rel.err={e:.2e}", flush=True)
        assert e <= 1e-13
    # alpha = 0 leaves y alone
    assert np.array_equal(_gpu(plan, 0.0, N.OP_NORMAL, x, y0, cuda), y0)


@pytest.mark.parametrize("form", ["uniform", "h2"])
def test_reduced_precision(cuda, mats, form):
    um = mats["sphere"][form]
    x = random_vector(um.ncols, 5)
    yr = np.zeros(um.nrows)
    oracle.uniform_mul_vec(um.rowcb, um.colcb, um.blocks, um.arena, 1.0, N.OP_NORMAL, x, yr)
    for prec, eps, bar in (("fp32", 0.0, 2e-6), ("valr", 1e-6, 1e-6)):
        plan = UPlan(um, precision=prec, eps=eps)
        y = _gpu(plan, 1.0, N.OP_NORMAL, x, np.zeros(um.nrows), cuda)
        e = rel(y, yr)
        print(f"[gpu] uplan sphere {form} {prec} rel.err={e:.2e}", flush=True)
        assert e <= bar
        plan.close()
