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
    respect to x and all factors against central finite differences of the
    same loss evaluated in FP64 by an FP64 plan of the stored values
    (h = 1e-3, 32 random directions, relative bar 1e-3), and the factor
    gradients of 100 sampled tiles against their closed forms, within the
    componentwise rounding bound of their FP32 panel sums
Each prints its measured errors."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.autograd import arena_tensor, hmatvec
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, MATERN_CUBE, MATERN_TLR, random_vector
from hlr_amd.plan import Plan
from tests.helpers import rel

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
    n = hm.n
    plan = Plan.from_hmatrix(hm, precision="fp32")
    print(f"[gpu] fullsize TLR n={n} tiles={len(hm.blocks)}", flush=True)
    F = N.FMT_F32
    U0 = arena_tensor(plan, F)
    Ua = U0.clone().requires_grad_(True)
    x = torch.as_tensor(random_vector(n, 41, gaussian=True), device=cuda).requires_grad_(True)
    b = torch.as_tensor(random_vector(n, 42, gaussian=True), device=cuda)

    def loss(xv, ua):
        r = hmatvec(xv, plan, {F: ua}) - b
        return 0.5 * torch.dot(r, r)

    L = loss(x, Ua)
    L.backward()
    torch.cuda.synchronize()
    with torch.no_grad():
        r = hmatvec(x.detach(), plan, {F: U0}) - b
    assert plan.arena_elems(N.FMT_F64) == 0

    # the stored factors in the leaves' own layout (exact: FP32 values in
    # FP64), through the plan's layout map
    src32 = torch.as_tensor(plan.arena_source(F), device=cuda)
    live32 = src32 >= 0
    deq = torch.zeros(hm.arena.size, dtype=torch.float64, device=cuda)
    deq[src32[live32]] = U0[live32].double()
    deq_h = deq.cpu().numpy()

    gref = np.zeros(n)
    oracle.mul_vec(hm.blocks, deq_h, 1.0, N.OP_TRANSPOSED, r.cpu().numpy(), gref)
    eg = rel(x.grad.cpu().numpy(), gref)
    print(f"[gpu] TLR n={n} dL/dx vs oracle transposed rel.err={eg:.3e}", flush=True)
    assert eg <= 2e-6

    # factor gradients in the leaves' layout
    g_gen = torch.zeros_like(deq)
    g_gen[src32[live32]] = Ua.grad[live32].double()
    del src32, live32

    # closed forms on 100 sampled tiles: dU = r_b (V^T x_b)^T, dV = x_b (U^T r_b)^T,
    # dD = r_b x_b^T.  On the GPU t = V^T x and s = U^T r are sums of at most
    # 64 FP32 products of FP32-rounded operands per panel (FP64 across
    # panels), so componentwise |t_gpu - t| <= g |V|^T |x| with g = 66 u,
    # u = 2^-24 (the standard bound for n-term floating point sums), and the
    # gradient passes through the FP32 arena dtype once more (u relative)
    u32 = 2.0 ** -24
    g = 66 * u32
    g_h = g_gen.cpu().numpy()
    x_h = x.detach().cpu().numpy()
    r_h = r.cpu().numpy()
    rng = np.random.default_rng(3)
    worst = 0.0
    for bi in rng.choice(len(hm.blocks), size=100, replace=False):
        blk = hm.blocks[bi]
        m, nn, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
        ro, co, oa, ob = int(blk["row_ofs"]), int(blk["col_ofs"]), int(blk["off_a"]), int(blk["off_b"])
        rb, xb = r_h[ro:ro + m], x_h[co:co + nn]
        if int(blk["kind"]) == N.BLOCK_DENSE:
            gD = g_h[oa:oa + m * nn].reshape(nn, m).T
            ref = np.outer(rb, xb)
            # one FP32 rounding of an FP64 product (itself within 2^-53)
            bound = u32 * (1 + 1e-6) * np.abs(ref) + 1e-300
            q = np.abs(gD - ref) / bound
        elif k:
            U = deq_h[oa:oa + m * k].reshape(k, m).T
            V = deq_h[ob:ob + nn * k].reshape(k, nn).T
            gU = g_h[oa:oa + m * k].reshape(k, m).T
            gV = g_h[ob:ob + nn * k].reshape(k, nn).T
            refU = np.outer(rb, V.T @ xb)
            refV = np.outer(xb, U.T @ rb)
            bU = np.outer(np.abs(rb), g * (np.abs(V).T @ np.abs(xb))) + u32 * (1 + 1e-6) * np.abs(refU) + 1e-300
            bV = np.outer(np.abs(xb), g * (np.abs(U).T @ np.abs(rb))) + u32 * (1 + 1e-6) * np.abs(refV) + 1e-300
            q = max((np.abs(gU - refU) / bU).max(), (np.abs(gV - refV) / bV).max())
        else:
            continue
        q = float(np.max(q))
        worst = max(worst, q)
        assert q <= 1.0, (int(bi), q)
    print(f"[gpu] TLR n={n} closed-form tile gradients over 100 tiles: worst error / rounding bound = {worst:.3f}",
          flush=True)
    del g_h, deq_h

    # central finite differences in FP64: an FP64 plan holding exactly the
    # stored values evaluates the loss, with the factors moved along a
    # direction given in the leaves' layout; h = 1e-3, 32 random directions
    # in (x, factors)
    p64 = Plan(hm.blocks, deq.cpu().numpy(), hm.nrows, hm.ncols, precision="fp64")
    src64 = torch.as_tensor(p64.arena_source(N.FMT_F64), device=cuda)
    live64 = src64 >= 0
    base64 = torch.zeros(p64.arena_elems(N.FMT_F64), dtype=torch.float64, device=cuda)
    base64[live64] = deq[src64[live64]]

    def loss64(xv, a64):
        p64.set_arena(N.FMT_F64, a64)
        y = torch.zeros(n, dtype=torch.float64, device=cuda)
        p64.apply(1.0, N.OP_NORMAL, xv, y)
        rr = y - b
        return 0.5 * float(torch.dot(rr, rr))

    gen = torch.Generator(device=cuda).manual_seed(11)
    h = 1e-3
    worst = 0.0
    xd = x.detach()
    with torch.no_grad():
        for _ in range(32):
            dx = torch.randn(n, generator=gen, dtype=torch.float64, device=cuda)
            dg = torch.randn(deq.numel(), generator=gen, dtype=torch.float64, device=cuda) * deq.abs()
            d64 = torch.zeros_like(base64)
            d64[live64] = dg[src64[live64]]
            lp = loss64(xd + h * dx, base64 + h * d64)
            lm = loss64(xd - h * dx, base64 - h * d64)
            fd = (lp - lm) / (2 * h)
            ax = float(torch.dot(x.grad, dx))
            au = float(torch.dot(g_gen, dg))
            dev = abs(fd - (ax + au))
            worst = max(worst, dev / (abs(ax) + abs(au)))
            assert dev <= 1e-3 * (abs(ax) + abs(au)), (fd, ax, au)
            del dg, d64
    print(f"[gpu] TLR n={n} worst FD relative deviation {worst:.2e} (FP64 loss, no allowance)", flush=True)
