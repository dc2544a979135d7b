"""Parity of the HIP product against the oracle (GPU).

Bars: plain FP64 path <= 1e-13 relative to the oracle's FP64 mul_vec on the
same H-matrix (forward, transposed, adjoint); FP32-accumulating storage
formats <= 2e-6 relative to the oracle run on the plan's own dequantised
factors; VALR storage <= eps relative to the FP64 oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, LOG1D, MATERN_CUBE, MATERN_TLR, random_vector
from hlr_amd.plan import Plan
from tests.helpers import dequantized_arena, rel

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-13
TOL_FP32ACC = 2e-6


@pytest.fixture(scope="module")
def log1d():
    return HMatrix(LOG1D)


@pytest.fixture(scope="module")
def sphere16k():
    return HMatrix(INVR_SPHERE.with_(n=16384))


def _gpu_apply(plan, alpha, op, x, y0, cuda, dtype=None):
    import torch
    dt = dtype or torch.float64
    xt = torch.as_tensor(x, dtype=dt, device=cuda)
    yt = torch.as_tensor(y0, dtype=dt, device=cuda).clone()
    print(f"[gpu] apply op={op} alpha={alpha} n={x.shape} prec={plan.precision}", flush=True)
    plan.apply(alpha, op, xt, yt)
    torch.cuda.synchronize()
    return yt.double().cpu().numpy()


def _oracle(hm, alpha, op, x, y0, arena=None):
    y = np.array(y0, dtype=np.float64, copy=True)
    oracle.mul_vec(hm.blocks, hm.arena if arena is None else arena, alpha, op, x, y)
    return y


@pytest.mark.parametrize("op", [N.OP_NORMAL, N.OP_TRANSPOSED, N.OP_ADJOINT, N.OP_CONJUGATE])
def test_fp64_log1d(cuda, log1d, op):
    hm = log1d
    plan = Plan.from_hmatrix(hm, precision="fp64")
    x = random_vector(hm.n, 2)
    y0 = random_vector(hm.n, 3)
    y = _gpu_apply(plan, 2.5, op, x, y0, cuda)
    yr = _oracle(hm, 2.5, op, x, y0)
    assert rel(y, yr) <= TOL_FP64


@pytest.mark.parametrize("op", [N.OP_NORMAL, N.OP_TRANSPOSED])
def test_fp64_sphere(cuda, sphere16k, op):
    hm = sphere16k
    plan = Plan.from_hmatrix(hm, precision="fp64")
    x = random_vector(hm.n, 4)
    y0 = np.zeros(hm.n)
    y = _gpu_apply(plan, 1.0, op, x, y0, cuda)
    yr = _oracle(hm, 1.0, op, x, y0)
    assert rel(y, yr) <= TOL_FP64


def test_alpha_zero_and_accumulate(cuda, log1d):
    import torch
    hm = log1d
    plan = Plan.from_hmatrix(hm, precision="fp64")
    x = torch.as_tensor(random_vector(hm.n, 5), device=cuda)
    y0 = torch.as_tensor(random_vector(hm.n, 6), device=cuda)
    y = y0.clone()
    plan.apply(0.0, N.OP_NORMAL, x, y)
    torch.cuda.synchronize()
    assert torch.equal(y, y0)                     # alpha = 0 leaves y untouched
    y1 = torch.zeros_like(y0)
    plan.apply(1.0, N.OP_NORMAL, x, y1)
    y2 = y1.clone()
    plan.apply(1.0, N.OP_NORMAL, x, y2)           # accumulates, never overwrites
    torch.cuda.synchronize()
    assert rel(y2.cpu().numpy(), (2 * y1).cpu().numpy()) <= 1e-15


def test_deterministic(cuda, sphere16k):
    import torch
    hm = sphere16k
    plan = Plan.from_hmatrix(hm, precision="fp64")
    x = torch.as_tensor(random_vector(hm.n, 7), device=cuda)
    outs = []
    for _ in range(3):
        y = torch.zeros_like(x)
        plan.apply(1.0, N.OP_NORMAL, x, y)
        plan.apply(1.0, N.OP_TRANSPOSED, x, y)
        outs.append(y)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[1], outs[2])


def test_multi_rhs(cuda, log1d):
    import torch
    hm = log1d
    plan = Plan.from_hmatrix(hm, precision="fp64")
    nrhs = 5
    X = np.stack([random_vector(hm.n, 10 + r) for r in range(nrhs)])
    Y0 = np.stack([random_vector(hm.n, 20 + r) for r in range(nrhs)])
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        Xt = torch.as_tensor(X, device=cuda)
        Yt = torch.as_tensor(Y0, device=cuda).clone()
        plan.apply(-1.5, op, Xt, Yt)
        torch.cuda.synchronize()
        Yr = Y0.copy()
        oracle.mul_vec(hm.blocks, hm.arena, -1.5, op, X, Yr)
        assert rel(Yt.cpu().numpy(), Yr) <= TOL_FP64


@pytest.mark.parametrize("nrhs", [2, 16, 18, 32, 33, 64, 70])
@pytest.mark.parametrize("prec", ["fp64", "valr"])
def test_multi_rhs_matrix_cores(cuda, sphere16k, prec, nrhs):
    """Forward products with several right-hand sides run on the matrix
    cores (16 x 16 x 4 MFMA tiles, rhs padded to 16, > 64 split)."""
    import torch
    hm = sphere16k
    plan = Plan.from_hmatrix(hm, precision=prec, eps=1e-6)
    arena = hm.arena if prec == "fp64" else dequantized_arena(plan, hm)
    X = np.stack([random_vector(hm.n, 100 + r) for r in range(nrhs)])
    Y0 = np.stack([random_vector(hm.n, 200 + r) for r in range(nrhs)])
    Xt = torch.as_tensor(X, device=cuda)
    Yt = torch.as_tensor(Y0, device=cuda).clone()
    print(f"[gpu] mrhs prec={prec} nrhs={nrhs}", flush=True)
    plan.apply(0.5, N.OP_NORMAL, Xt, Yt)
    torch.cuda.synchronize()
    Yr = Y0.copy()
    oracle.mul_vec(hm.blocks, arena, 0.5, N.OP_NORMAL, X, Yr)
    got = Yt.cpu().numpy()
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32ACC
    for r in range(nrhs):
        assert rel(got[r] - Y0[r], Yr[r] - Y0[r]) <= tol, r


@pytest.mark.parametrize("nrhs", [16, 17])
def test_multi_rhs_workspace_alignment(cuda, sphere16k, nrhs):
    """A caller workspace that is only 8-byte aligned takes the partial sums
    one right-hand side per thread (pairs need 16-byte loads); both give
    bitwise the product computed with the library's own workspace, as the
    summation order is the same."""
    import torch
    hm = sphere16k
    plan = Plan.from_hmatrix(hm, precision="valr", eps=1e-6)
    X = torch.as_tensor(np.stack([random_vector(hm.n, 300 + r) for r in range(nrhs)]), device=cuda)
    ref = torch.zeros_like(X)
    print(f"[gpu] mrhs workspace alignment nrhs={nrhs}", flush=True)
    plan.apply(1.0, N.OP_NORMAL, X, ref)
    nd = plan.workspace_bytes(N.OP_NORMAL, nrhs) // 8 + 1
    buf = torch.empty(nd + 1, dtype=torch.float64, device=cuda)
    for off in (0, 1):                      # 16-byte aligned, then 8-byte aligned
        y = torch.zeros_like(X)
        plan.apply_ws(1.0, N.OP_NORMAL, X, y, buf[off:off + nd])
        torch.cuda.synchronize()
        assert torch.equal(y, ref), off
        assert float(y.abs().max()) > 0


def test_fp32_vectors(cuda, log1d):
    import torch
    hm = log1d
    plan = Plan.from_hmatrix(hm, precision="fp64")
    x = random_vector(hm.n, 8).astype(np.float32)
    y = _gpu_apply(plan, 1.0, N.OP_NORMAL, x, np.zeros(hm.n, np.float32), cuda, dtype=torch.float32)
    yr = _oracle(hm, 1.0, N.OP_NORMAL, x.astype(np.float64), np.zeros(hm.n))
    assert rel(y, yr) <= 1e-6


@pytest.mark.parametrize("prec,eps,force", [("fp32", 0.0, None), ("valr", 1e-4, None), ("valr", 1e-6, None),
                                             ("valr", 1e-6, "fp8_e4m3"), ("valr", 1e-6, "bf16")])
@pytest.mark.parametrize("op", [N.OP_NORMAL, N.OP_TRANSPOSED])
def test_lowprec_vs_dequantized(cuda, sphere16k, prec, eps, force, op):
    """The kernels must reproduce the oracle on the stored (dequantised)
    factors up to FP32 accumulation.  Forcing every low-rank column into
    FP8-E4M3 (or BF16) makes the hardware decode carry most of the result,
    so a decode that differed from the OCP table would fail here."""
    hm = sphere16k
    plan = Plan.from_hmatrix(hm, precision=prec, eps=eps, force_format=force)
    info = plan.info()
    if force is not None:
        f = ("fp64", "fp32", "bf16", "fp8_e4m3").index(force)
        assert info["cols_by_format"][f] == sum(info["cols_by_format"])
    deq = dequantized_arena(plan, hm)
    x = random_vector(hm.n, 9)
    y = _gpu_apply(plan, 1.0, op, x, np.zeros(hm.n), cuda)
    yr = _oracle(hm, 1.0, op, x, np.zeros(hm.n), arena=deq)
    assert rel(y, yr) <= TOL_FP32ACC


@pytest.mark.parametrize("eps", [1e-4, 1e-6])
def test_valr_error_vs_fp64(cuda, sphere16k, eps):
    """Mixed-precision storage chosen against eps stays within eps of the
    FP64 oracle on the same H-matrix."""
    hm = HMatrix(INVR_SPHERE.with_(n=16384, eps=eps))
    plan = Plan.from_hmatrix(hm, precision="valr", eps=eps)
    x = random_vector(hm.n, 11)
    y = _gpu_apply(plan, 1.0, N.OP_NORMAL, x, np.zeros(hm.n), cuda)
    yr = _oracle(hm, 1.0, N.OP_NORMAL, x, np.zeros(hm.n))
    assert rel(y, yr) <= eps


def test_tlr(cuda):
    import torch
    hm = HMatrix(MATERN_TLR.with_(n=4096))
    plan = Plan.from_hmatrix(hm, precision="fp64")
    assert plan.info()["one_level"] == 1
    x = random_vector(hm.n, 12)
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        xt = torch.as_tensor(x, device=cuda)
        yt = torch.zeros_like(xt)
        plan.tlr_apply(1.0, op, xt, yt)
        torch.cuda.synchronize()
        yr = np.zeros(hm.n)
        oracle.tlr_mul_vec(hm.blocks, hm.arena, 1.0, op, x, yr, nthreads=4)
        assert rel(yt.cpu().numpy(), yr) <= TOL_FP64


def test_tlr_rejects_h(cuda, log1d):
    import torch
    plan = Plan.from_hmatrix(log1d, precision="fp64")
    x = torch.zeros(log1d.n, dtype=torch.float64, device=cuda)
    with pytest.raises(N.HmvmError):
        plan.tlr_apply(1.0, N.OP_NORMAL, x, x.clone())
