"""Packed storage formats on the GPU: block floating point (B8, B16, B25)
and adaptive float (A12, A16, A24).

Bars: the product of a packed plan is within 2e-6 of the oracle run on the
plan's own dequantised factors (FP32 accumulation inside a panel), for every
width forced on all low-rank columns, forward and transposed, for several
right-hand sides, and for the adjoint; the plan chosen against eps is within
eps of the FP64 oracle on the same H-matrix (AFL: within eps where its
widest code, 18 mantissa bits, and FP32 / FP64 carry the columns).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, MATERN_TLR, random_vector
from hlr_amd.plan import FORMAT_NAMES, Plan
from tests.helpers import dequantized_arena, rel

pytestmark = pytest.mark.gpu

TOL_FP32ACC = 2e-6
FAMILY = {"bfp": (N.FMT_B8, N.FMT_B25), "afl": (N.FMT_A12, N.FMT_A24)}
WIDTHS = [("bfp", f) for f in FORMAT_NAMES[N.FMT_B8:N.FMT_B25 + 1]] + \
         [("afl", f) for f in FORMAT_NAMES[N.FMT_A12:N.FMT_A24 + 1]]


@pytest.fixture(scope="module")
def sphere16k():
    return HMatrix(INVR_SPHERE.with_(n=16384))


def _apply(plan, alpha, op, x, y0, cuda):
    import torch
    xt = torch.as_tensor(x, device=cuda)
    yt = torch.as_tensor(y0, device=cuda).clone()
    print(f"[gpu] packed apply op={op} shape={x.shape}", flush=True)
    plan.apply(alpha, op, xt, yt)
    torch.cuda.synchronize()
    return yt.cpu().numpy()


@pytest.mark.parametrize("prec,force", WIDTHS)
@pytest.mark.parametrize("op", [N.OP_NORMAL, N.OP_TRANSPOSED])
def test_forced_width_vs_dequantized(cuda, sphere16k, prec, force, op):
    """Every low-rank column in one width: the width's decode carries most
    of the result, so a field offset, sign or exponent handling that differed
    from the host packing would fail here."""
    hm = sphere16k
    plan = Plan.from_hmatrix(hm, precision=prec, eps=1e-6, force_format=force)
    c = plan.info()["cols_by_format"]
    assert c[FORMAT_NAMES.index(force)] == sum(c)
    deq = dequantized_arena(plan, hm)
    x = random_vector(hm.n, 21)
    y0 = random_vector(hm.n, 22)
    y = _apply(plan, -0.75, op, x, y0, cuda)
    yr = y0.copy()
    oracle.mul_vec(hm.blocks, deq, -0.75, op, x, yr)
    assert rel(y - y0, yr - y0) <= TOL_FP32ACC


@pytest.mark.parametrize("prec", ["bfp", "afl"])
@pytest.mark.parametrize("eps", [1e-4, 1e-6, 1e-8])
def test_error_vs_fp64(cuda, prec, eps):
    """Widths chosen against eps keep the product within eps of the FP64
    oracle (forward and transposed)."""
    hm = HMatrix(INVR_SPHERE.with_(n=16384, eps=eps))
    plan = Plan.from_hmatrix(hm, precision=prec, eps=eps)
    c = plan.info()["cols_by_format"]
    first, last = FAMILY[prec]
    assert sum(c[first:last + 1]) > 0.5 * sum(c)
    x = random_vector(hm.n, 23)
    deq = dequantized_arena(plan, hm)
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        y = _apply(plan, 1.0, op, x, np.zeros(hm.n), cuda)
        yr = np.zeros(hm.n)
        oracle.mul_vec(hm.blocks, hm.arena, 1.0, op, x, yr)
        assert rel(y, yr) <= eps, (op, rel(y, yr))
        yq = np.zeros(hm.n)
        oracle.mul_vec(hm.blocks, deq, 1.0, op, x, yq)
        assert rel(y, yq) <= TOL_FP32ACC


@pytest.mark.parametrize("prec", ["bfp", "afl"])
@pytest.mark.parametrize("nrhs", [2, 16, 32, 33, 64])
def test_multi_rhs(cuda, sphere16k, prec, nrhs):
    """Several right-hand sides on the matrix cores with all of the family's
    widths in one plan (the V pass splits a width into accumulator sets of
    columns)."""
    import torch
    hm = sphere16k
    plan = Plan.from_hmatrix(hm, precision=prec, eps=1e-6)
    c = plan.info()["cols_by_format"]
    first, last = FAMILY[prec]
    assert all(c[f] > 0 for f in range(first, last + 1)), c
    deq = dequantized_arena(plan, hm)
    X = np.stack([random_vector(hm.n, 300 + r) for r in range(nrhs)])
    Y0 = np.stack([random_vector(hm.n, 400 + r) for r in range(nrhs)])
    Xt = torch.as_tensor(X, device=cuda)
    Yt = torch.as_tensor(Y0, device=cuda).clone()
    print(f"[gpu] {prec} mrhs nrhs={nrhs}", flush=True)
    plan.apply(0.5, N.OP_NORMAL, Xt, Yt)
    torch.cuda.synchronize()
    Yr = Y0.copy()
    oracle.mul_vec(hm.blocks, deq, 0.5, N.OP_NORMAL, X, Yr)
    got = Yt.cpu().numpy()
    for r in range(nrhs):
        assert rel(got[r] - Y0[r], Yr[r] - Y0[r]) <= TOL_FP32ACC, r


@pytest.mark.parametrize("prec", ["bfp", "afl"])
def test_grad(cuda, prec):
    """Adjoint of a packed plan: A^T r against the oracle's transposed
    product on the stored factors, factor gradients against their closed
    forms (V^T x and U^T r accumulate in FP32 inside a panel)."""
    import torch
    hm = HMatrix(MATERN_TLR.with_(n=4096))
    plan = Plan.from_hmatrix(hm, precision=prec, eps=1e-6)
    n = hm.n
    x = torch.as_tensor(random_vector(n, 41, gaussian=True), device=cuda)
    r = torch.as_tensor(random_vector(n, 42, gaussian=True), device=cuda)
    gx = torch.zeros_like(x)
    print(f"[gpu] {prec} grad n={n}", flush=True)
    garena = plan.grad(x, r, gx)
    torch.cuda.synchronize()
    deq = dequantized_arena(plan, hm)
    r_h, x_h = r.cpu().numpy(), x.cpu().numpy()
    gx_ref = np.zeros(n)
    oracle.mul_vec(hm.blocks, deq, 1.0, N.OP_TRANSPOSED, r_h, gx_ref)
    assert rel(gx.cpu().numpy(), gx_ref) <= TOL_FP32ACC
    g_host = [g.cpu().numpy() if g is not None else np.zeros(0) for g in garena]
    ar = [plan.arena_numpy(f) for f in range(N.NFMT)]
    rng = np.random.default_rng(1)
    for bi in rng.choice(len(hm.blocks), size=min(40, len(hm.blocks)), replace=False):
        blk = hm.blocks[bi]
        m, nn, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
        ro, co = int(blk["row_ofs"]), int(blk["col_ofs"])
        gA, gB = plan.unpack_block(int(bi), g_host, as_f64=True)
        if int(blk["kind"]) == N.BLOCK_DENSE:
            assert rel(gA, np.outer(r_h[ro:ro + m], x_h[co:co + nn])) <= 1e-12
        elif k:
            U, V = plan.unpack_block(int(bi), ar)
            assert rel(gA, np.outer(r_h[ro:ro + m], V.T @ x_h[co:co + nn])) <= TOL_FP32ACC
            assert rel(gB, np.outer(x_h[co:co + nn], U.T @ r_h[ro:ro + m])) <= TOL_FP32ACC
