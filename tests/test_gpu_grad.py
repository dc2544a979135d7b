"""Adjoint of the H-MVM (GPU): hmvm_grad against the oracle's transposed
mul_vec, closed-form factor gradients, and central finite differences.

Loss L = 1/2 ||A x - b||^2, r = A x - b:
  dL/dx   = A^T r                       (oracle transposed mul_vec)
  dL/dU_b = r_b (S_b V_b^T x_b)^T       dL/dV_b = x_b (S_b U_b^T r_b)^T
  dL/dD_b = r_b x_b^T
Finite differences: central, h = 1e-3 (FP32 storage) or 1e-5 (FP64), random
directions in factor space.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, MATERN_TLR, MATERN_CUBE, random_vector
from hlr_amd.plan import Plan
from tests.helpers import dequantized_arena, rel

pytestmark = pytest.mark.gpu


def _loss(plan, x, b):
    import torch
    y = torch.zeros_like(b)
    plan.apply(1.0, N.OP_NORMAL, x, y)
    r = y - b
    return 0.5 * float(torch.dot(r, r)), r


@pytest.mark.parametrize("case", ["tlr_fp32", "h_fp64"])
def test_grad(cuda, case):
    import torch
    if case == "tlr_fp32":
        hm = HMatrix(MATERN_TLR.with_(n=4096))
        plan = Plan.from_hmatrix(hm, precision="fp32")
    else:
        hm = HMatrix(MATERN_CUBE.with_(n=4096))
        plan = Plan.from_hmatrix(hm, precision="fp64")
    n = hm.n
    x = torch.as_tensor(random_vector(n, 31, gaussian=True), device=cuda)
    b = torch.as_tensor(random_vector(n, 32, gaussian=True), device=cuda)
    L0, r = _loss(plan, x, b)
    gx = torch.zeros_like(x)
    print(f"[gpu] grad case={case} n={n}", flush=True)
    garena = plan.grad(x, r, gx)
    torch.cuda.synchronize()

    # dL/dx against the oracle's transposed product on the stored factors
    deq = dequantized_arena(plan, hm)
    r_h = r.cpu().numpy()
    x_h = x.cpu().numpy()
    gx_ref = np.zeros(n)
    oracle.mul_vec(hm.blocks, deq, 1.0, N.OP_TRANSPOSED, r_h, gx_ref)
    # FP32 factors accumulate in FP32 inside a panel (2e-6); FP64 plans 1e-12
    assert rel(gx.cpu().numpy(), gx_ref) <= (2e-6 if case == "tlr_fp32" else 1e-12)

    # per-block closed forms
    g_host = [g.cpu().numpy() if g is not None else np.zeros(0) for g in garena]
    rng = np.random.default_rng(0)
    for bi in rng.choice(len(hm.blocks), size=min(40, len(hm.blocks)), replace=False):
        blk = hm.blocks[bi]
        m, nn, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
        ro, co = int(blk["row_ofs"]), int(blk["col_ofs"])
        gA, gB = plan.unpack_block(int(bi), g_host, as_f64=True)
        if int(blk["kind"]) == N.BLOCK_DENSE:
            ref = np.outer(r_h[ro:ro + m], x_h[co:co + nn])
            assert rel(gA, ref) <= 1e-12
        elif k:
            U, V = plan.unpack_block(int(bi), [plan.arena_numpy(f) for f in range(N.NFMT)])
            t = V.T @ x_h[co:co + nn]
            s = U.T @ r_h[ro:ro + m]
            # t and s are accumulated in FP32 inside a panel for FP32 factors
            tol = 2e-6 if case == "tlr_fp32" else 1e-10
            assert rel(gA, np.outer(r_h[ro:ro + m], t)) <= tol
            assert rel(gB, np.outer(x_h[co:co + nn], s)) <= tol

    # central finite differences along random factor-space directions
    fmt = 1 if case == "tlr_fp32" else 0
    tdt = torch.float32 if fmt == 1 else torch.float64
    base = torch.as_tensor(plan.arena_numpy(fmt), device=cuda)
    g = garena[fmt]
    # L is a polynomial of degree 6 along a direction that moves U and V
    # together, so central differences carry a truncation error of about
    # h^2 / 6 L''' (about 1e-6 relative at h = 1e-3): the FP64 check takes
    # h = 1e-5, below its 1e-6 bar; FP32 storage rounds at 6e-8, so h = 1e-3
    h = 1e-3 if fmt == 1 else 1e-5
    gen = torch.Generator(device="cpu").manual_seed(5)
    for _ in range(4):
        d = (torch.randn(base.numel(), generator=gen, dtype=torch.float64).to(cuda) * base.double().abs())
        plan.set_arena(fmt, (base.double() + h * d).to(tdt).contiguous())
        Lp, _ = _loss(plan, x, b)
        plan.set_arena(fmt, (base.double() - h * d).to(tdt).contiguous())
        Lm, _ = _loss(plan, x, b)
        fd = (Lp - Lm) / (2 * h)
        an = float(torch.dot(g, d))
        tol = 2e-3 if fmt == 1 else 1e-6
        assert abs(fd - an) <= tol * max(abs(an), 1e-30), (fd, an)
    plan.set_arena(fmt, base)
