"""torch.autograd binding over the C-ABI (GPU): the tile-low-rank Matern
problem with FP32 factors, L = 1/2 ||A x - b||^2, gradients with respect to
x and to every factor checked against central finite differences (h = 1e-3,
32 random directions) and against the oracle's transposed mul_vec."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.autograd import arena_tensor, hmatvec
from hlr_amd.hmatrix import HMatrix, MATERN_TLR, random_vector
from hlr_amd.plan import Plan
from tests.helpers import dequantized_arena, rel

pytestmark = pytest.mark.gpu


def test_tlr_fp32_autograd(cuda):
    import torch
    hm = HMatrix(MATERN_TLR.with_(n=8192))
    plan = Plan.from_hmatrix(hm, precision="fp32")
    print(f"[gpu] autograd TLR n={hm.n} tiles={len(hm.blocks)}", flush=True)
    F = N.FMT_F32
    U0 = arena_tensor(plan, F)
    Ua = U0.clone().requires_grad_(True)
    x = torch.as_tensor(random_vector(hm.n, 41, gaussian=True), device=cuda).requires_grad_(True)
    b = torch.as_tensor(random_vector(hm.n, 42, gaussian=True), device=cuda)

    def loss(xv, ua):
        r = hmatvec(xv, plan, {F: ua}) - b
        return 0.5 * torch.dot(r, r)

    L = loss(x, Ua)
    L.backward()
    torch.cuda.synchronize()

    # dL/dx = A^T r against the oracle on the stored factors
    with torch.no_grad():
        r = hmatvec(x.detach(), plan, {F: U0}) - b
    deq = dequantized_arena(plan, hm)
    gref = np.zeros(hm.n)
    oracle.mul_vec(hm.blocks, deq, 1.0, N.OP_TRANSPOSED, r.cpu().numpy(), gref)
    assert rel(x.grad.cpu().numpy(), gref) <= 2e-6

    # central finite differences, h = 1e-3, 32 random directions in (x, factors)
    gen = torch.Generator(device="cpu").manual_seed(11)
    h = 1e-3
    worst = 0.0
    with torch.no_grad():
        for _ in range(32):
            dx = torch.randn(hm.n, generator=gen, dtype=torch.float64).to(cuda)
            du = (torch.randn(U0.numel(), generator=gen, dtype=torch.float64).to(cuda) * U0.double().abs())
            lp = float(loss(x + h * dx, (U0.double() + h * du).float()))
            lm = float(loss(x - h * dx, (U0.double() - h * du).float()))
            fd = (lp - lm) / (2 * h)
            ax = float(torch.dot(x.grad, dx))
            au = float(torch.dot(Ua.grad.double(), du))
            # relative to the size of the two contributions (their sum may cancel)
            worst = max(worst, abs(fd - (ax + au)) / (abs(ax) + abs(au)))
    print(f"[gpu] autograd worst FD relative deviation {worst:.2e}", flush=True)
    assert worst <= 1e-3
