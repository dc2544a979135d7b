"""Generator (construction side) checks on the CPU.

Kernels against closed forms and library routines (scipy Bessel K, numerical
quadrature), the dense linear algebra against numpy, ACA + truncation
against the kernel block, clustering and admissibility invariants.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, LOG1D, MATERN_CUBE, MATERN_TLR, Problem, random_vector


def _p(a):
    return C.c_void_p(a.ctypes.data)


def test_qr_svd_vs_numpy():
    lib = N.hlrgen()
    rng = np.random.default_rng(0)
    for m, k in ((50, 7), (200, 31), (8, 8)):
        A = np.asfortranarray(rng.standard_normal((m, k)))
        Q = np.zeros((m, k), order="F")
        R = np.zeros((k, k), order="F")
        assert lib.hlrgen_qr(m, k, _p(A), _p(Q), _p(R)) == 0
        assert np.allclose(Q @ R, A, atol=1e-13 * np.abs(A).max())
        assert np.allclose(Q.T @ Q, np.eye(k), atol=1e-13)
        assert np.allclose(np.tril(R, -1), 0)
    for k in (1, 5, 40):
        A = np.asfortranarray(rng.standard_normal((k, k)) @ np.diag(np.logspace(0, -12, k)))
        U = np.zeros((k, k), order="F"); V = np.zeros((k, k), order="F"); S = np.zeros(k)
        assert lib.hlrgen_svd(k, _p(A), _p(U), _p(S), _p(V)) == 0
        assert np.allclose(S, np.linalg.svd(A, compute_uv=False), rtol=1e-10, atol=1e-15)
        assert np.allclose((U * S) @ V.T, A, atol=1e-13)
        assert np.all(np.diff(S) <= 0)


def test_truncate():
    lib = N.hlrgen()
    rng = np.random.default_rng(1)
    m, n, r = 120, 90, 20
    sv = np.logspace(0, -10, r)
    U = np.asfortranarray(np.linalg.qr(rng.standard_normal((m, r)))[0] * sv)
    V = np.asfortranarray(np.linalg.qr(rng.standard_normal((n, r)))[0])
    Uo = np.zeros((m, r), order="F"); Vo = np.zeros((n, r), order="F"); So = np.zeros(r)
    k = C.c_int64()
    for eps in (1e-3, 1e-6):
        assert lib.hlrgen_truncate(m, n, r, _p(U), _p(V), eps, _p(Uo), _p(So), _p(Vo), C.byref(k)) == 0
        kk = k.value
        assert kk == int(np.sum(sv > eps * sv[0]))
        M = U @ V.T
        Mk = (Uo[:, :kk] * So[:kk]) @ Vo[:, :kk].T
        assert np.linalg.norm(M - Mk, 2) <= 1.01 * sv[kk] + 1e-15
        assert np.allclose(So[:kk], sv[:kk], rtol=1e-10)


def test_aca_block():
    hm = HMatrix(MATERN_CUBE.with_(n=4096))
    lib = N.hlrgen()
    # an admissible block of the generated tree: take the largest low-rank leaf
    b = hm.blocks[np.argmax(np.where(hm.blocks["kind"] != 0, hm.blocks["nrows"] * hm.blocks["ncols"], 0))]
    rows = np.arange(b["row_ofs"], b["row_ofs"] + b["nrows"], dtype=np.int64)
    cols = np.arange(b["col_ofs"], b["col_ofs"] + b["ncols"], dtype=np.int64)
    kmax = min(len(rows), len(cols))
    U = np.zeros((len(rows), kmax), order="F"); V = np.zeros((len(cols), kmax), order="F")
    k = C.c_int64()
    assert lib.hlrgen_aca(hm._h, len(rows), _p(rows), len(cols), _p(cols), 1e-6, kmax, _p(U), _p(V),
                          C.byref(k)) == 0
    K = hm.eval(rows, cols)
    kk = k.value
    err = np.linalg.norm(K - U[:, :kk] @ V[:, :kk].T) / np.linalg.norm(K)
    assert kk < kmax and err <= 1e-5


def test_matern_closed_form_vs_bessel():
    from scipy.special import gamma, kv
    hm = HMatrix(MATERN_CUBE.with_(n=512))
    X = hm.coords
    rng = np.random.default_rng(2)
    i = rng.integers(0, hm.n, 60)
    j = rng.integers(0, hm.n, 60)
    vals = hm.eval(i, j)
    ell, nu = 0.1, 1.5
    for a in range(60):
        for b in range(0, 60, 7):
            r = np.linalg.norm(X[i[a]] - X[j[b]])
            if r == 0:
                ref = 1.0
            else:
                s = math.sqrt(2 * nu) * r / ell
                ref = 2 ** (1 - nu) / gamma(nu) * s ** nu * kv(nu, s)
            assert vals[a, b] == pytest.approx(ref, rel=1e-12, abs=1e-300)


def test_log_cell_average_and_galerkin():
    from scipy.integrate import dblquad
    n = 64
    h = 1.0 / n
    hm = HMatrix(LOG1D.with_(n=n, ntile=16))
    d = hm.eval([0], [0])[0, 0]
    avg, _ = dblquad(lambda s, t: math.log(abs(s - t)) if s != t else 0.0, 0, h, 0, h, epsabs=1e-13)
    assert d == pytest.approx(avg / h ** 2, rel=1e-6)
    assert d == pytest.approx(math.log(h) - 1.5, rel=1e-14)
    g = HMatrix(LOG1D.with_(n=n, ntile=16, kernel=N.KERNEL_LOG_GALERKIN, geometry=N.GEOM_GRID1D))
    inv = np.argsort(g.perm)              # external -> internal
    for a, b in ((0, 0), (3, 4), (10, 2), (0, 63)):
        v = g.eval([inv[a]], [inv[b]])[0, 0]
        ref, _ = dblquad(lambda s, t: math.log(abs(s - t)) if s != t else 0.0,
                         a * h, (a + 1) * h, b * h, (b + 1) * h, epsabs=1e-14)
        assert v == pytest.approx(-ref, rel=1e-6)


def test_invr_nn_diagonal():
    hm = HMatrix(INVR_SPHERE.with_(n=2000))
    X = hm.coords
    D = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2)
    np.fill_diagonal(D, np.inf)
    nn = np.sort(D, axis=1)[:, :8]
    ref = np.mean(1.0 / nn, axis=1)
    got = np.array([hm.eval([i], [i])[0, 0] for i in range(hm.n)])
    assert np.allclose(got, ref, rtol=1e-12)
    # points are on the unit sphere
    assert np.allclose(np.linalg.norm(X, axis=1), 1.0)


def test_cluster_invariants():
    hm = HMatrix(MATERN_CUBE.with_(n=5000))
    assert sorted(hm.perm.tolist()) == list(range(hm.n))
    b = hm.blocks
    # leaves tile the matrix exactly once
    assert int((b["nrows"].astype(np.int64) * b["ncols"]).sum()) == hm.n * hm.n
    dense = b["kind"] == N.BLOCK_DENSE
    assert (b["nrows"][dense] <= 65).all() and (b["ncols"][dense] <= 65).all()
    t = HMatrix(MATERN_TLR.with_(n=2048, ntile=256))
    assert set(t.blocks["nrows"].tolist()) == {256}
    assert len(t.blocks) == 64


def test_seeded_reproducible():
    a = HMatrix(INVR_SPHERE.with_(n=3000))
    b = HMatrix(INVR_SPHERE.with_(n=3000))
    assert np.array_equal(a.blocks, b.blocks) and np.array_equal(a.arena, b.arena)
    assert np.array_equal(random_vector(100, 5), random_vector(100, 5))
    x = random_vector(10000, 5)
    assert x.min() >= -1 and x.max() <= 1 and abs(x.mean()) < 0.05
