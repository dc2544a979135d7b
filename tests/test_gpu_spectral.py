"""Spectral norm by power iteration on the device (GPU), the reference's
generic consumer of mul_vec (arith/detail/norm.hh:448-527).

The start vector is the reference's: std::default_random_engine
(minstd_rand0, seed 1) through uniform_real_distribution<double>(-1, 1),
restated here after libstdc++'s generate_canonical (two draws per double),
so a numpy replay of the same iteration on the dense expansion must give the
same lambda after a fixed number of steps; run to convergence, lambda must
equal numpy's 2-norm of the expansion."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd.hmatrix import HMatrix, LOG1D, MATERN_CUBE
from hlr_amd.plan import Plan

pytestmark = pytest.mark.gpu


def reference_start_vector(n: int) -> np.ndarray:
    m, a = 2 ** 31 - 1, 16807           # minstd_rand0, default seed 1
    state = 1
    r = float(m - 1)                    # max - min + 1 = 2^31 - 2
    x = np.empty(n)
    for i in range(n):
        s, tmp = 0.0, 1.0
        for _ in range(2):              # ceil(53 / floor(log2 r)) = 2 draws
            state = (a * state) % m
            s += float(state - 1) * tmp
            tmp *= r
        u = s / tmp
        x[i] = u * 2.0 - 1.0
    return x / np.linalg.norm(x)


def numpy_power(A: np.ndarray, x: np.ndarray, steps: int, squared: bool) -> float:
    lam = 1.0
    for _ in range(steps):
        if squared:
            y = A.T @ (A @ x)
            lam = np.sqrt(abs(x @ y))
        else:
            y = A @ x
            lam = np.linalg.norm(y)
        x = y / np.linalg.norm(y)
    return float(lam)


@pytest.mark.parametrize("problem", ["log", "matern"])
def test_spectral_norm(cuda, problem):
    prob = LOG1D.with_(n=1536) if problem == "log" else MATERN_CUBE.with_(n=2048)
    hm = HMatrix(prob)
    A = oracle.dense_expand(hm.blocks, hm.arena, hm.n, hm.n)
    p = Plan.from_hmatrix(hm, precision="fp64")
    x0 = reference_start_vector(hm.n)
    # fixed number of steps (stopping rules disabled by a vanishing tolerance)
    lam, its = p.spectral_norm(atol=1e-300, max_it=12)
    ref = numpy_power(A, x0, 12, squared=True)
    print(f"[gpu] spectral {problem} n={hm.n}: 12 steps lambda={lam:.15e} numpy={ref:.15e}", flush=True)
    assert its == 12
    assert abs(lam - ref) <= 1e-12 * ref
    # converged: the 2-norm of the matrix
    lam, its = p.spectral_norm(atol=1e-14, max_it=5000)
    two = float(np.linalg.norm(A, 2))
    print(f"[gpu] spectral {problem} n={hm.n}: converged in {its} steps lambda={lam:.15e} 2-norm={two:.15e}",
          flush=True)
    assert abs(lam - two) <= 1e-9 * two
    # symmetric kernel: |lambda| of the unsquared iteration is the 2-norm as well
    lam1, _ = p.spectral_norm(atol=1e-14, max_it=20000, squared=False)
    assert abs(lam1 - two) <= 1e-6 * two
    # default stopping rule (sqrt(eps)) stops early with a close value
    lam2, its2 = p.spectral_norm()
    assert its2 <= max(5, hm.n // 10) and abs(lam2 - two) <= 1e-6 * two
