"""A plan built from an H-matrix file computes bitwise the same products as
the plan built from the same leaves in memory (GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, random_vector
from hlr_amd.plan import Plan

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "bfp"])
def test_plan_from_file_products(cuda, tmp_path, prec):
    import torch
    hm = HMatrix(INVR_SPHERE.with_(n=16384))
    path = tmp_path / "sphere.hm"
    hm.save(path)
    a = Plan.from_hmatrix(hm, precision=prec, eps=1e-6)
    b = Plan.from_file(path, precision=prec, eps=1e-6)
    x = torch.as_tensor(random_vector(hm.n, 5), device=cuda)
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        ya = torch.zeros_like(x)
        yb = torch.zeros_like(x)
        print(f"[gpu] plan from file prec={prec} op={op} n={hm.n}", flush=True)
        a.apply(1.5, op, x, ya)
        b.apply(1.5, op, x, yb)
        torch.cuda.synchronize()
        assert torch.equal(ya, yb)
        assert np.linalg.norm(ya.cpu().numpy()) > 0
