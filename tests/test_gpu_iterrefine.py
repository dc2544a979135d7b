"""Linear iteration x <- x - W (A x - b) and the inversion error |I - A W|_2
on the device (GPU), the reference's iterative refinement
(programs/iterrefine.hh:20-80) and hlr::norm::inv_error_2
(arith/norm.hh:115-137).

A is the FP64 plan of a Matern covariance (SPD, condition about 1.6e4 at
n = 2048); W is an approximate inverse in A's block structure (exact dense
leaves of A^-1, truncated SVDs of its admissible blocks), optionally damped
by omega, stored in FP64 or in a compressed precision -- the reference runs
the same iteration with an FP64 and an FP32 H-LU inverse as W.

Bars: the step count equals a numpy replay of the iteration on the dense
expansions of A and of W's stored (dequantised) factors, and the per-step
|r|, |c|, |e| equal it to 1e-6 with an FP64 W, and within the propagated
effect of FP32 accumulation inside W's panels with a compressed W;
|I - A W| after a fixed number of steps equals numpy's power iteration from
the reference's start vector, and converged it equals the 2-norm of the
dense I - A W."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd.hmatrix import HMatrix, MATERN_CUBE, random_vector
from hlr_amd.plan import Plan, inv_error_2, linear_iteration
from tests.helpers import dequantized_arena, inverse_blocks
from tests.test_gpu_spectral import numpy_power, reference_start_vector

pytestmark = pytest.mark.gpu

# per-step norms with an FP64 W agree with the dense replay up to the
# cancellation in r = A x - b (|A| |x| eps / |r|, about 1e-7 near 1e-8)
TOL_FP64_HIST = 1e-6
# a compressed W accumulates W r in FP32 inside a panel: relative error of
# the product at most this (the bar of the block floating point and mixed
# precision product tests)
TOL_FP32ACC = 2e-6


def propagated_tolerances(A, Wq, ref, delta=TOL_FP32ACC, safety=4.0):
    """Bounds on |history - replay| when every product W r carries a
    relative error delta: the error of x obeys dx_{i+1} = (I - W A) dx_i +
    delta |c_i|, so |dx_{i+1}| <= rho |dx_i| + delta |c_i| with rho =
    |I - W A|_2; then |r_i| moves by at most |A| |dx_i|, |c_i| by
    delta |c_i| + |W A| |dx_i|, and |e_i| (taken after the step) by
    |dx_{i+1}|."""
    WA = Wq @ A
    rho = np.linalg.norm(np.eye(A.shape[0]) - WA, 2)
    nA, nWA = np.linalg.norm(A, 2), np.linalg.norm(WA, 2)
    steps = len(ref)
    dx = np.zeros(steps + 1)
    for i in range(steps):
        dx[i + 1] = rho * dx[i] + delta * ref[i, 1]
    tol = np.stack([nA * dx[:steps], delta * ref[:, 1] + nWA * dx[:steps], dx[1:]], axis=1)
    return safety * tol + 1e-12 * ref + 1e-14


@pytest.fixture(scope="module")
def system():
    hm = HMatrix(MATERN_CUBE.with_(n=2048))
    A = oracle.dense_expand(hm.blocks, hm.arena, hm.n, hm.n)
    sol = random_vector(hm.n, 1)
    sol /= np.linalg.norm(sol)
    return hm, A, sol, A @ sol


def replay(A, W, b, sol, tol=1e-8, max_steps=100, residual_stop=False):
    """The iteration of programs/iterrefine.hh:43-81 in numpy, from x = 0
    with r = A x - b."""
    x = np.zeros_like(b)
    r = A @ x - b
    nb = np.linalg.norm(b)
    hist = []
    if residual_stop and np.linalg.norm(r) <= tol * nb:
        return np.zeros((0, 3)), x
    for _ in range(max_steps):
        c = W @ r
        x = x - c
        e = np.linalg.norm(sol - x)
        hist.append((np.linalg.norm(r), np.linalg.norm(c), e))
        r = A @ x - b
        if (np.linalg.norm(r) <= tol * nb) if residual_stop else (e <= tol):
            break
    return np.array(hist), x


def _plans(hm, Wh, prec):
    Ap = Plan.from_hmatrix(hm, precision="fp64")
    Wp = Plan(Wh.blocks, Wh.arena, hm.n, hm.n, precision=prec, eps=1e-6)
    Wq = oracle.dense_expand(Wh.blocks, dequantized_arena(Wp, Wh), hm.n, hm.n)
    return Ap, Wp, Wq


@pytest.mark.parametrize("omega", [1.0, 0.5])
@pytest.mark.parametrize("prec", ["fp64", "valr", "bfp"])
def test_linear_iteration(cuda, system, prec, omega):
    import torch
    hm, A, sol, b = system
    Wh = inverse_blocks(hm, A, 1e-5, omega)
    Ap, Wp, Wq = _plans(hm, Wh, prec)
    ref, xr = replay(A, Wq, b, sol)
    bt = torch.as_tensor(b, device=cuda)
    st = torch.as_tensor(sol, device=cuda)
    xt = torch.zeros(hm.n, dtype=torch.float64, device=cuda)
    print(f"[gpu] linear_iteration prec={prec} omega={omega} n={hm.n}", flush=True)
    steps, hist = linear_iteration(Ap, Wp, bt, xt, sol=st)
    print(f"[gpu]   {steps} steps (replay {len(ref)}), |e| = {hist[-1, 2]:.3e}", flush=True)
    assert steps == len(ref)
    assert hist[-1, 2] <= 1e-8 < hist[-2, 2]
    tol = TOL_FP64_HIST * ref + 1e-13 if prec == "fp64" else propagated_tolerances(A, Wq, ref)
    assert np.all(np.abs(hist - ref) <= tol), np.max(np.abs(hist - ref) / tol)
    assert np.linalg.norm(xt.cpu().numpy() - xr) <= 1e-9
    # the refinement reaches the FP64 solution whatever W's storage
    assert np.linalg.norm(xt.cpu().numpy() - sol) <= 1e-8


def test_linear_iteration_residual_stop(cuda, system):
    """Without a known solution the iteration stops on |A x - b| <= tol |b|
    and reports no |e|; a start value that already solves the system takes
    no step."""
    import torch
    hm, A, sol, b = system
    Wh = inverse_blocks(hm, A, 1e-5, 0.5)
    Ap, Wp, Wq = _plans(hm, Wh, "bfp")
    ref, _ = replay(A, Wq, b, sol, tol=1e-10, residual_stop=True)
    bt = torch.as_tensor(b, device=cuda)
    xt = torch.zeros(hm.n, dtype=torch.float64, device=cuda)
    print(f"[gpu] linear_iteration residual stop n={hm.n}", flush=True)
    steps, hist = linear_iteration(Ap, Wp, bt, xt, tol=1e-10, max_steps=200)
    assert steps == len(ref) and steps > 20
    assert np.all(np.isnan(hist[:, 2]))
    tol = propagated_tolerances(A, Wq, ref)
    assert np.all(np.abs(hist[:, :2] - ref[:, :2]) <= tol[:, :2])
    r = A @ xt.cpu().numpy() - b
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b) * 1.001
    # step limit: stops after max_steps with the history of those steps
    xt.zero_()
    steps, hist = linear_iteration(Ap, Wp, bt, xt, tol=1e-10, max_steps=5)
    assert steps == 5 and hist.shape == (5, 3)
    assert np.all(np.abs(hist[:, :2] - ref[:5, :2]) <= tol[:5, :2])
    # already solved: no step, x untouched
    xt = torch.as_tensor(np.linalg.solve(A, b), device=cuda)
    x0 = xt.clone()
    steps, hist = linear_iteration(Ap, Wp, bt, xt, tol=1e-6)
    assert steps == 0 and hist.shape == (0, 3) and torch.equal(xt, x0)


@pytest.mark.parametrize("prec", ["fp64", "bfp"])
def test_inv_error_2(cuda, system, prec):
    hm, A, _, _ = system
    Wh = inverse_blocks(hm, A, 1e-5)
    Ap, Wp, Wq = _plans(hm, Wh, prec)
    E = np.eye(hm.n) - A @ Wq
    x0 = reference_start_vector(hm.n)
    print(f"[gpu] inv_error_2 prec={prec} n={hm.n}", flush=True)
    err, its = inv_error_2(Ap, Wp, atol=1e-300, max_it=12)
    ref = numpy_power(E, x0, 12, squared=True)
    assert its == 12
    # t = x - A (W x) cancels to |t| = 0.018 |x| while |W x| reaches
    # |A^-1| |x| = 470 |x|: the FP32 accumulation error of W x (about 1e-7
    # |W x|) shows in t amplified by up to |A| |A^-1| / 0.018
    rtol = 1e-10 if prec == "fp64" else 1e-2
    assert abs(err - ref) <= rtol * ref
    err, its = inv_error_2(Ap, Wp, atol=1e-14, max_it=2000)
    two = float(np.linalg.norm(E, 2))
    print(f"[gpu]   converged in {its} steps: {err:.15e}, 2-norm {two:.15e}", flush=True)
    assert abs(err - two) <= max(rtol, 1e-9) * two
