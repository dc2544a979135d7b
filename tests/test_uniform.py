"""Uniform-H and H^2 matrices on the CPU: the generator's conversion, the
oracle's product (arith/uniform.hh, arith/h2.hh restated) and host-only
device plans.

Pins of the oracle:
  - its dense expansion equals an expansion assembled here in numpy
    (explicit bases, transfers multiplied out son by son, couplings);
  - its product equals numpy's product with that expansion for all ops,
    alpha and accumulation (1e-13);
  - an H^2 matrix and the uniform-H matrix with its expanded bases and the
    same couplings give the same product (1e-13).
The conversion keeps the H-matrix to within a small multiple of eps.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, LOG1D, MATERN_CUBE, random_vector
from hlr_amd.uniform import UniformHMatrix, UPlan, build_uniform
from tests.helpers import rel


@pytest.fixture(scope="module")
def mats():
    out = {}
    for name, prob in (("log1d", LOG1D.with_(n=2048)), ("sphere", INVR_SPHERE.with_(n=2048)),
                       ("matern", MATERN_CUBE.with_(n=2048))):
        hm = HMatrix(prob)
        out[name] = (hm, build_uniform(hm, 1e-6, nested=False), build_uniform(hm, 1e-6, nested=True))
    return out


def _basis(um: UniformHMatrix, cb, i):
    """expanded basis of cluster i, assembled here"""
    c = cb[i]
    s, k = int(c["size"]), int(c["rank"])
    if c["off_basis"] >= 0:
        return um.arena[c["off_basis"]:c["off_basis"] + s * k].reshape(k, s).T
    out = np.zeros((s, k))
    for j in np.nonzero(cb["parent"] == i)[0]:
        sj, kj = int(cb[j]["size"]), int(cb[j]["rank"])
        if kj == 0 or cb[j]["off_transfer"] < 0:
            continue
        E = um.arena[cb[j]["off_transfer"]:cb[j]["off_transfer"] + kj * k].reshape(k, kj).T
        o = int(cb[j]["ofs"]) - int(c["ofs"])
        out[o:o + sj] += _basis(um, cb, j) @ E
    return out


def _expand(um: UniformHMatrix):
    A = np.zeros((um.nrows, um.ncols))
    for b in um.blocks:
        r0, c0, m, n = int(b["row_ofs"]), int(b["col_ofs"]), int(b["nrows"]), int(b["ncols"])
        if int(b["kind"]) == N.BLOCK_DENSE:
            A[r0:r0 + m, c0:c0 + n] += um.arena[b["off"]:b["off"] + m * n].reshape(n, m).T
            continue
        kr, kc = int(um.rowcb[b["rowcb"]]["rank"]), int(um.colcb[b["colcb"]]["rank"])
        if kr == 0 or kc == 0:
            continue
        S = um.arena[b["off"]:b["off"] + kr * kc].reshape(kc, kr).T
        A[r0:r0 + m, c0:c0 + n] += _basis(um, um.rowcb, int(b["rowcb"])) @ S @ _basis(um, um.colcb, int(b["colcb"])).T
    return A


@pytest.mark.parametrize("name", ["log1d", "sphere", "matern"])
def test_oracle_expansion_and_product(mats, name):
    hm, uni, h2 = mats[name]
    assert not uni.nested and h2.nested
    assert (h2.colcb["off_transfer"] >= 0).any() and (h2.rowcb["off_transfer"] >= 0).any()
    for um in (uni, h2):
        A = _expand(um)
        Ao = oracle.uniform_dense_expand(um.rowcb, um.colcb, um.blocks, um.arena, um.nrows, um.ncols)
        assert rel(Ao, A) <= 1e-13
        x = random_vector(um.ncols, 5)
        y0 = random_vector(um.nrows, 6)
        for op, M in ((N.OP_NORMAL, A), (N.OP_TRANSPOSED, A.T), (N.OP_ADJOINT, A.T), (N.OP_CONJUGATE, A)):
            y = y0.copy()
            oracle.uniform_mul_vec(um.rowcb, um.colcb, um.blocks, um.arena, -1.5, op, x, y)
            assert rel(y, y0 - 1.5 * (M @ x)) <= 1e-13, (name, um.nested, op)
        y = y0.copy()
        oracle.uniform_mul_vec(um.rowcb, um.colcb, um.blocks, um.arena, 0.0, N.OP_NORMAL, x, y)
        assert np.array_equal(y, y0)


@pytest.mark.parametrize("name", ["log1d", "matern"])
def test_h2_equals_uniform_with_expanded_bases(mats, name):
    _, _, h2 = mats[name]
    # the same matrix with every basis explicit (expanded) and no transfers
    parts = [h2.arena]
    pos = h2.arena.size
    cbs = []
    for cb in (h2.rowcb, h2.colcb):
        new = cb.copy()
        new["parent"] = -1
        new["off_transfer"] = -1
        for i in range(len(cb)):
            if int(cb[i]["rank"]) == 0:
                continue
            B = _basis(h2, cb, i)
            new[i]["off_basis"] = pos
            parts.append(B.T.reshape(-1))
            pos += B.size
        cbs.append(new)
    flat = UniformHMatrix(cbs[0], cbs[1], h2.blocks, np.concatenate(parts), h2.nrows, h2.ncols, False)
    x = random_vector(h2.ncols, 7)
    for op in (N.OP_NORMAL, N.OP_TRANSPOSED):
        ya = np.zeros(h2.nrows)
        yb = np.zeros(h2.nrows)
        oracle.uniform_mul_vec(h2.rowcb, h2.colcb, h2.blocks, h2.arena, 1.0, op, x, ya)
        oracle.uniform_mul_vec(flat.rowcb, flat.colcb, flat.blocks, flat.arena, 1.0, op, x, yb)
        assert rel(ya, yb) <= 1e-13


@pytest.mark.parametrize("name", ["log1d", "sphere", "matern"])
def test_conversion_accuracy(mats, name):
    hm, uni, h2 = mats[name]
    AH = oracle.dense_expand(hm.blocks, hm.arena, hm.nrows, hm.ncols)
    for um in (uni, h2):
        A = oracle.uniform_dense_expand(um.rowcb, um.colcb, um.blocks, um.arena, um.nrows, um.ncols)
        e = np.linalg.norm(A - AH) / np.linalg.norm(AH)
        print(f"{name} nested={um.nested}: |A_uni - A_H| / |A_H| = {e:.2e}")
        assert e <= 2e-5


def test_host_plan_info_and_checks(mats):
    _, uni, h2 = mats["matern"]
    for um in (uni, h2):
        p = UPlan(um, device=-1)
        info = p.info()
        assert info["nested"] == int(um.nested)
        assert info["ncoupling"] == int((um.blocks["kind"] == N.BLOCK_COUPLING).sum())
        assert info["ndense"] == int((um.blocks["kind"] == N.BLOCK_DENSE).sum())
        assert (info["levels"] > 0) == um.nested
        p.close()
    # damaged records are rejected
    bad = uni.blocks.copy()
    i = int(np.nonzero(bad["kind"] == N.BLOCK_COUPLING)[0][0])
    bad[i]["nrows"] += 1
    with pytest.raises(N.HmvmError, match="BAD_BLOCK"):
        UPlan(UniformHMatrix(uni.rowcb, uni.colcb, bad, uni.arena, uni.nrows, uni.ncols, False), device=-1)
    cyc = h2.colcb.copy()
    root = int(np.nonzero(cyc["parent"] == -1)[0][0])
    child = int(np.nonzero(cyc["parent"] == root)[0][0])
    cyc[root]["parent"] = child
    with pytest.raises(N.HmvmError, match="BAD_BLOCK"):
        UPlan(UniformHMatrix(h2.rowcb, cyc, h2.blocks, h2.arena, h2.nrows, h2.ncols, True), device=-1)
    short = uni.arena[: uni.arena.size // 2]
    with pytest.raises(N.HmvmError, match="BAD_BLOCK"):
        UPlan(UniformHMatrix(uni.rowcb, uni.colcb, uni.blocks, short, uni.nrows, uni.ncols, False), device=-1)
