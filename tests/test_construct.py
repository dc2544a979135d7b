"""Construction split between host and device, checked on the CPU: the
generator's skip_lowrank mode (tree, blocks and dense leaves as in the full
build, low-rank leaves left at rank 0) and the argument checks of
hmvm_lowrank_build that run before any device work."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from hlr_amd import _native as N
from hlr_amd.hmatrix import LOG1D, MATERN_CUBE, HMatrix


def _skip(problem):
    lib = N.hlrgen()
    cfg = N.HlrgenConfig()
    lib.hlrgen_default_config(C.byref(cfg))
    for name, _ in N.HlrgenConfig._fields_:
        if name not in ("verbose", "skip_lowrank"):
            setattr(cfg, name, getattr(problem, name))
    cfg.skip_lowrank = 1
    h = C.c_void_p()
    N.check(lib.hlrgen_build(C.byref(cfg), C.byref(h)), "hlrgen_build")
    nb = int(lib.hlrgen_nblocks(h))
    bt = N.block_dtype()
    blocks = np.frombuffer((C.c_char * (nb * bt.itemsize)).from_address(lib.hlrgen_blocks(h)), dtype=bt).copy()
    alen = int(lib.hlrgen_arena_len(h))
    arena = np.ctypeslib.as_array((C.c_double * alen).from_address(lib.hlrgen_arena(h))).copy()
    lib.hlrgen_free(h)
    return blocks, arena


@pytest.mark.parametrize("prob", [LOG1D.with_(n=2048), MATERN_CUBE.with_(n=2048)])
def test_skip_lowrank_keeps_structure_and_dense_leaves(prob):
    hm = HMatrix(prob)
    blocks, arena = _skip(prob)
    for f in ("row_ofs", "col_ofs", "nrows", "ncols", "kind"):
        assert np.array_equal(blocks[f], hm.blocks[f]), f
    lr = blocks["kind"] != N.BLOCK_DENSE
    assert lr.any() and (blocks["rank"][lr] == 0).all()
    for b, hb in zip(blocks[~lr], hm.blocks[~lr]):
        m, n = int(b["nrows"]), int(b["ncols"])
        assert np.array_equal(arena[b["off_a"]:b["off_a"] + m * n], hm.arena[hb["off_a"]:hb["off_a"] + m * n])


def _fn(hm):
    coords = np.ascontiguousarray(hm.coords)
    perm = np.ascontiguousarray(hm.perm)
    fn = N.HmvmKernelFn(kind=hm.problem.kernel, n=hm.n, h=1.0 / hm.n, ell=hm.problem.matern_ell,
                        sigma=hm.problem.matern_sigma, coords=coords.ctypes.data, perm=perm.ctypes.data)
    return fn, (coords, perm)


def test_lowrank_build_rejects_bad_leaves_before_device_work():
    hm = HMatrix(LOG1D.with_(n=1024))
    fn, keep = _fn(hm)
    lib = N.hmvm()
    h = C.c_void_p()
    leaf = np.zeros(1, dtype=N.lr_leaf_dtype())
    # a leaf holding diagonal entries (overlapping row and column ranges)
    leaf[0] = (0, 32, 64, 64)
    assert lib.hmvm_lowrank_build(C.byref(fn), leaf.ctypes.data, 1, 1e-6, 1, 0, C.byref(h)) == 2   # BAD_BLOCK
    # outside the point set, empty
    leaf[0] = (1000, 0, 64, 64)
    assert lib.hmvm_lowrank_build(C.byref(fn), leaf.ctypes.data, 1, 1e-6, 1, 0, C.byref(h)) == 2
    leaf[0] = (512, 0, 0, 64)
    assert lib.hmvm_lowrank_build(C.byref(fn), leaf.ctypes.data, 1, 1e-6, 1, 0, C.byref(h)) == 2
    # bad eps and kernel
    leaf[0] = (512, 0, 64, 64)
    assert lib.hmvm_lowrank_build(C.byref(fn), leaf.ctypes.data, 1, 0.0, 1, 0, C.byref(h)) == 1  # INVALID_ARG
    bad = N.HmvmKernelFn(kind=7, n=hm.n, h=1.0 / hm.n, ell=0.1, sigma=1.0, coords=keep[0].ctypes.data, perm=None)
    assert lib.hmvm_lowrank_build(C.byref(bad), leaf.ctypes.data, 1, 1e-6, 1, 0, C.byref(h)) == 1
    gal = N.HmvmKernelFn(kind=N.KFN_LOG_GALERKIN, n=hm.n, h=1.0 / hm.n, ell=0.1, sigma=1.0,
                         coords=keep[0].ctypes.data, perm=None)
    assert lib.hmvm_lowrank_build(C.byref(gal), leaf.ctypes.data, 1, 1e-6, 1, 0, C.byref(h)) == 1
    assert not h.value
