"""Parity at the workloads' full sizes (GPU), against the oracle:
  - Matern nu = 1.5 on n = 65536 points in the unit cube, FP64 plan, normal
    and transposed products: relative error <= 1e-13
  - 1/r on n = 262144 points on the unit sphere, eps in {1e-4, 1e-6, 1e-8}:
    FP64 plan <= 1e-13, mixed-precision and block floating point plans
    <= eps (normal and transposed), and 32 right-hand sides at eps = 1e-6
    on the matrix cores
  - tile-low-rank Matern on n = 131072 (tile 512, FP32 factors): gradient
    of L = 1/2 ||A x - b||^2 with respect to x against the oracle's
    transposed product on the stored factors, and the gradients with
    respect to x and all factors against central finite differences
    (h = 1e-3, 32 random directions; the bar adds the measured rounding noise
    of the FP32-accumulated loss, divided by h)
Each prints its measured errors."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.autograd import arena_tensor, hmatvec
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, MATERN_CUBE, MATERN_TLR, random_vector
from hlr_amd.plan import Plan
from tests.helpers import dequantized_arena, rel

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def _gpu(plan, alpha, op, X, Y0, cuda):
    import torch
    x = torch.as_tensor(np.ascontiguousarray(X), device=cuda)
    y = torch.as_tensor(np.ascontiguousarray(Y0), device=cuda)
    plan.apply(alpha, op, x, y)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _ref(hm, alpha, op, X, Y0):
    Y = np.array(Y0, copy=True)
    oracle.mul_vec(hm.blocks, hm.arena, alpha, op, np.ascontiguousarray(X), Y)
    return Y


def test_matern_cube_fp64(cuda):
    hm = HMatrix(MATERN_CUBE)
    print(f"[gpu] fullsize matern cube n={hm.n} leaves={len(hm.blocks)}", flush=True)
    p = Plan.from_hmatrix(hm, precision="fp64")
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        x = random_vector(hm.n, 3 + op)
        y0 = random_vector(hm.n, 30 + op)
        e = rel(_gpu(p, 2.0, op, x, y0, cuda), _ref(hm, 2.0, op, x, y0))
        print(f"[gpu] matern cube n={hm.n} fp64 op={op} rel.err={e:.3e}", flush=True)
        assert e <= 1e-13


@pytest.mark.parametrize("eps", [1e-4, 1e-6, 1e-8])
def test_sphere_eps_sweep(cuda, eps):
    hm = HMatrix(INVR_SPHERE.with_(eps=eps))
    print(f"[gpu] fullsize sphere n={hm.n} eps={eps:g} leaves={len(hm.blocks)}", flush=True)
    p64 = Plan.from_hmatrix(hm, precision="fp64")
    pmp = Plan.from_hmatrix(hm, precision="valr", eps=eps)
    pbf = Plan.from_hmatrix(hm, precision="bfp", eps=eps)
    info, ibf = pmp.info(), pbf.info()
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        x = random_vector(hm.n, 2 + op)
        y0 = np.zeros(hm.n)
        yr = _ref(hm, 1.0, op, x, y0)
        e64 = rel(_gpu(p64, 1.0, op, x, y0, cuda), yr)
        emp = rel(_gpu(pmp, 1.0, op, x, y0, cuda), yr)
        ebf = rel(_gpu(pbf, 1.0, op, x, y0, cuda), yr)
        print(f"[gpu] sphere n={hm.n} eps={eps:g} op={op} fp64 rel.err={e64:.3e} mixed rel.err={emp:.3e} "
              f"bfp rel.err={ebf:.3e} cols/format mixed={info['cols_by_format']} bfp={ibf['cols_by_format']} "
              f"bytes fp64={p64.info()['datasize']:.4g} mixed={info['datasize']:.4g} bfp={ibf['datasize']:.4g}",
              flush=True)
        assert e64 <= 1e-13
        assert emp <= eps
        assert ebf <= eps


def test_sphere_32_rhs(cuda):
    eps = 1e-6
    hm = HMatrix(INVR_SPHERE.with_(eps=eps))
    pmp = Plan.from_hmatrix(hm, precision="valr", eps=eps)
    X = np.stack([random_vector(hm.n, 100 + j) for j in range(32)])
    Y0 = np.zeros_like(X)
    Yr = _ref(hm, 1.0, N.OP_NORMAL, X, Y0)
    Y = _gpu(pmp, 1.0, N.OP_NORMAL, X, Y0, cuda)
    errs = [rel(Y[j], Yr[j]) for j in range(32)]
    print(f"[gpu] sphere n={hm.n} eps={eps:g} nrhs=32 mixed max rel.err={max(errs):.3e}", flush=True)
    assert max(errs) <= eps


def test_tlr_gradients(cuda):
    import torch
    hm = HMatrix(MATERN_TLR)
    plan = Plan.from_hmatrix(hm, precision="fp32")
    print(f"[gpu] fullsize TLR n={hm.n} tiles={len(hm.blocks)}", flush=True)
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
    with torch.no_grad():
        r = hmatvec(x.detach(), plan, {F: U0}) - b
    deq = dequantized_arena(plan, hm)
    gref = np.zeros(hm.n)
    oracle.mul_vec(hm.blocks, deq, 1.0, N.OP_TRANSPOSED, r.cpu().numpy(), gref)
    eg = rel(x.grad.cpu().numpy(), gref)
    print(f"[gpu] TLR n={hm.n} dL/dx vs oracle transposed rel.err={eg:.3e}", flush=True)
    assert eg <= 2e-6
    # rounding noise of one loss evaluation: the GPU loss (FP32 accumulation
    # inside tiles) against the same loss in FP64 on the stored factors
    xh = x.detach().cpu().numpy()
    ax_ref = np.zeros(hm.n)
    oracle.mul_vec(hm.blocks, deq, 1.0, N.OP_NORMAL, xh, ax_ref)
    del deq
    L_ref = 0.5 * float(np.dot(ax_ref - b.cpu().numpy(), ax_ref - b.cpu().numpy()))
    dL = abs(float(L) - L_ref)
    print(f"[gpu] TLR n={hm.n} loss rounding |L_gpu - L_fp64| = {dL:.3e} (L = {L_ref:.6e})", flush=True)

    # central finite differences, h = 1e-3, 32 random directions in (x, factors);
    # directions drawn on the device (the factor arena has ~1.7e9 entries).
    # L is quadratic in (x, factors), so central differences have no
    # truncation error: what remains is the rounding of the two loss values,
    # at most about 2 dL / (2 h), allowed on top of the relative bar.
    gen = torch.Generator(device=cuda).manual_seed(11)
    h = 1e-3
    noise = 4.0 * dL / h
    worst = 0.0
    with torch.no_grad():
        Ud = U0.double()
        for _ in range(32):
            dx = torch.randn(hm.n, generator=gen, dtype=torch.float64, device=cuda)
            du = torch.randn(U0.numel(), generator=gen, dtype=torch.float64, device=cuda) * Ud.abs()
            lp = float(loss(x + h * dx, (Ud + h * du).float()))
            lm = float(loss(x - h * dx, (Ud - h * du).float()))
            fd = (lp - lm) / (2 * h)
            ax = float(torch.dot(x.grad, dx))
            au = float(torch.dot(Ua.grad.double(), du))
            dev = abs(fd - (ax + au))
            worst = max(worst, dev / (abs(ax) + abs(au)))
            assert dev <= 1e-3 * (abs(ax) + abs(au)) + noise, (fd, ax, au, noise)
            del du
    print(f"[gpu] TLR n={hm.n} worst FD relative deviation {worst:.2e} "
          f"(rounding allowance {noise:.2e} absolute)", flush=True)
