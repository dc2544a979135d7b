"""H-/TLR-matrices from the native generator (libhlrgen.so).

`HMatrix` holds the flattened leaf list (in hlr::mul_vec visiting order) and
the FP64 factor arena produced by hlrgen_build.  With construct="device" the
host builds the cluster tree, the block tree and the dense leaves, and the
low-rank leaves are built on the GPU (hmvm_lowrank_build: ACA and SVD
truncation, one workgroup per leaf).  The named problems restate the
workloads the product is measured on (kernel, geometry, sizes, seeds).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _native as N


@dataclass(frozen=True)
class Problem:
    """Parameters of one generated matrix (mirrors hlrgen_config)."""
    kernel: int = N.KERNEL_LOG
    geometry: int = N.GEOM_LINE
    n: int = 4096
    seed: int = 1
    ntile: int = 64
    format: int = N.FORMAT_H
    eta: float = 2.0
    eps: float = 1e-6
    matern_ell: float = 0.1
    matern_sigma: float = 1.0
    nn_diag: int = 8
    lrformat: int = N.LR_UV
    nthreads: int = 0
    tlr_std_adm: int = 0

    def with_(self, **kw) -> "Problem":
        return replace(self, **kw)


# named workloads -----------------------------------------------------------

#: 1D log kernel on n uniform random points in [0,1], leaf 64, eta 2, ACA 1e-6
LOG1D = Problem(kernel=N.KERNEL_LOG, geometry=N.GEOM_LINE, n=4096)
#: Matern nu=1.5, l=0.1 on n uniform points in the unit cube
MATERN_CUBE = Problem(kernel=N.KERNEL_MATERN, geometry=N.GEOM_CUBE, n=65536, matern_ell=0.1)
#: 1/r on n uniform points on the unit sphere, 8-neighbour diagonal
INVR_SPHERE = Problem(kernel=N.KERNEL_INVR, geometry=N.GEOM_SPHERE, n=262144, nn_diag=8)
#: Matern nu=1.5, l=0.1, tile-low-rank with tile 512
MATERN_TLR = Problem(kernel=N.KERNEL_MATERN, geometry=N.GEOM_CUBE, n=131072, ntile=512,
                     format=N.FORMAT_TLR)


def batch_problems(count: int = 64, seed: int = 7, nmin: int = 2048, nmax: int = 32768):
    """`count` Matern problems with l ~ U[0.05,0.3], n log-uniform in [nmin,nmax]
    and nrhs uniform in {1,4,16,64} (seeded); returns [(Problem, nrhs)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        ell = float(rng.uniform(0.05, 0.3))
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        nrhs = int(rng.choice([1, 4, 16, 64]))
        out.append((Problem(kernel=N.KERNEL_MATERN, geometry=N.GEOM_CUBE, n=n, matern_ell=ell,
                            seed=1000 + i), nrhs))
    return out


class HMatrix:
    """A generated H-matrix: leaves + FP64 arena (owned by the native side)."""

    def __init__(self, problem: Problem, verbose: bool = False, construct: str = "host", device: int = 0):
        if construct not in ("host", "device"):
            raise ValueError(f"construct must be 'host' or 'device', not {construct!r}")
        lib = N.hlrgen()
        cfg = N.HlrgenConfig()
        lib.hlrgen_default_config(C.byref(cfg))
        for name, _ in N.HlrgenConfig._fields_:
            if name in ("verbose", "skip_lowrank"):
                continue
            setattr(cfg, name, getattr(problem, name))
        cfg.verbose = 1 if verbose else 0
        cfg.skip_lowrank = 1 if construct == "device" else 0
        h = C.c_void_p()
        N.check(lib.hlrgen_build(C.byref(cfg), C.byref(h)), "hlrgen_build")
        self._h = h
        self.problem = problem
        self.n = int(lib.hlrgen_n(h))
        nb = int(lib.hlrgen_nblocks(h))
        bt = N.block_dtype()
        raw = (C.c_char * (nb * bt.itemsize)).from_address(lib.hlrgen_blocks(h)) if nb else b""
        self.blocks = np.frombuffer(raw, dtype=bt, count=nb).copy() if nb else np.zeros(0, bt)
        alen = int(lib.hlrgen_arena_len(h))
        self.arena = np.ctypeslib.as_array((C.c_double * alen).from_address(lib.hlrgen_arena(h))) \
            if alen else np.zeros(0)
        self.perm = np.ctypeslib.as_array((C.c_int64 * self.n).from_address(lib.hlrgen_perm(h))).copy()
        self.coords = np.ctypeslib.as_array(
            (C.c_double * (3 * self.n)).from_address(lib.hlrgen_coords(h))).reshape(self.n, 3).copy()
        self.build_seconds = float(lib.hlrgen_build_seconds(h))
        self.construct = construct
        self.device_seconds = None
        if construct == "device":
            self._build_lowrank_on_device(device)

    def _build_lowrank_on_device(self, device: int) -> None:
        """Low-rank leaves (left at rank 0 by the host) from hmvm_lowrank_build;
        the arena grows by their factors in the host generator's layout (U,
        V, S per leaf, each padded to 4 values; S folded into the factor
        with fewer rows for HLRGEN_LR_UV)."""
        p = self.problem
        lr = np.nonzero(self.blocks["kind"] != N.BLOCK_DENSE)[0]
        leaves = np.zeros(len(lr), dtype=N.lr_leaf_dtype())
        leaves["row0"] = self.blocks["row_ofs"][lr]
        leaves["col0"] = self.blocks["col_ofs"][lr]
        leaves["nrows"] = self.blocks["nrows"][lr]
        leaves["ncols"] = self.blocks["ncols"][lr]
        coords = np.ascontiguousarray(self.coords)
        perm = np.ascontiguousarray(self.perm)
        fn = N.HmvmKernelFn(kind=p.kernel, n=self.n, h=1.0 / self.n, ell=p.matern_ell, sigma=p.matern_sigma,
                            coords=coords.ctypes.data, perm=perm.ctypes.data)
        lib = N.hmvm()
        h = C.c_void_p()
        N.check(lib.hmvm_lowrank_build(C.byref(fn), leaves.ctypes.data if len(lr) else None, len(lr), p.eps, 1,
                                       device, C.byref(h)), "hmvm_lowrank_build")
        try:
            ranks = np.zeros(len(lr), dtype=np.int32)
            N.check(lib.hmvm_lowrank_ranks(h, ranks.ctypes.data), "hmvm_lowrank_ranks")
            m = self.blocks["nrows"][lr].astype(np.int64)
            n = self.blocks["ncols"][lr].astype(np.int64)
            k = ranks.astype(np.int64)
            pad4 = lambda v: (v + 3) & ~np.int64(3)
            sizes = np.stack([pad4(m * k), pad4(n * k), pad4(k)], axis=1)
            base = self.arena.size
            ends = base + np.cumsum(sizes.reshape(-1))
            starts = (ends - sizes.reshape(-1)).reshape(-1, 3)
            arena = np.zeros(int(ends[-1]) if len(lr) else base)
            arena[:base] = self.arena
            off_a = np.ascontiguousarray(starts[:, 0])
            off_b = np.ascontiguousarray(starts[:, 1])
            off_s = np.ascontiguousarray(starts[:, 2])
            fold = 1 if p.lrformat == N.LR_UV else 0
            N.check(lib.hmvm_lowrank_copy(h, off_a.ctypes.data, off_b.ctypes.data, off_s.ctypes.data, fold,
                                          arena.ctypes.data, arena.size), "hmvm_lowrank_copy")
            self.device_seconds = float(lib.hmvm_lowrank_seconds(h))
            ks, cs = C.c_double(), C.c_double()
            N.check(lib.hmvm_lowrank_timing(h, C.byref(ks), C.byref(cs)), "hmvm_lowrank_timing")
            self.device_kernel_seconds, self.device_copy_seconds = float(ks.value), float(cs.value)
        finally:
            lib.hmvm_lowrank_free(h)
        self.blocks["rank"][lr] = ranks
        self.blocks["off_a"][lr] = off_a
        self.blocks["off_b"][lr] = off_b
        self.blocks["off_s"][lr] = off_s
        self.arena = arena

    # the arena view is only valid while the native matrix lives
    def close(self) -> None:
        if getattr(self, "_h", None):
            N.hlrgen().hlrgen_free(self._h)
            self._h = None
            self.arena = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nrows(self) -> int:
        return self.n

    @property
    def ncols(self) -> int:
        return self.n

    def eval(self, rows, cols) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        out = np.zeros((len(rows), len(cols)), order="F")
        N.check(N.hlrgen().hlrgen_eval(self._h, len(rows), rows.ctypes.data, len(cols), cols.ctypes.data,
                                       out.ctypes.data), "hlrgen_eval")
        return out

    def save(self, path) -> None:
        """Store the leaves and the FP64 arena in an H-matrix file."""
        save_hmatrix(path, self.blocks, self.arena, self.nrows, self.ncols)

    def stats(self) -> dict:
        b = self.blocks
        dense = b["kind"] == N.BLOCK_DENSE
        lr = ~dense
        m = b["nrows"].astype(np.float64)
        n = b["ncols"].astype(np.float64)
        k = b["rank"].astype(np.float64)
        return {
            "n": self.n, "leaves": int(len(b)), "dense": int(dense.sum()), "lowrank": int(lr.sum()),
            "kmax": int(k[lr].max()) if lr.any() else 0,
            "kmean": float(k[lr].mean()) if lr.any() else 0.0,
            "dense_bytes_fp64": float((m * n)[dense].sum() * 8),
            "lowrank_bytes_fp64": float((k * (m + n))[lr].sum() * 8),
            "build_seconds": self.build_seconds,
        }


def random_vector(n: int, seed: int, lo: float = -1.0, hi: float = 1.0, gaussian: bool = False) -> np.ndarray:
    """Seeded std::mt19937_64 vector: U[lo,hi] or N(0,1)."""
    out = np.empty(n)
    N.hlrgen().hlrgen_random_vector(seed, n, lo, hi, 1 if gaussian else 0, out.ctypes.data)
    return out


class StoredHMatrix:
    """An H-matrix read back from an H-matrix file: leaves + FP64 arena
    (numpy-owned), with the attributes of HMatrix that plans and the oracle
    use (blocks, arena, n, nrows, ncols)."""

    def __init__(self, blocks: np.ndarray, arena: np.ndarray, nrows: int, ncols: int):
        self.blocks, self.arena = blocks, arena
        self.nrows, self.ncols = nrows, ncols
        self.n = nrows

    def save(self, path) -> None:
        save_hmatrix(path, self.blocks, self.arena, self.nrows, self.ncols)


def save_hmatrix(path, blocks: np.ndarray, arena: np.ndarray, nrows: int, ncols: int) -> None:
    """Write leaves + FP64 arena as an H-matrix file (hmvm_hmatrix_write)."""
    blocks = np.ascontiguousarray(blocks, dtype=N.block_dtype())
    arena = np.ascontiguousarray(arena, dtype=np.float64)
    N.check(N.hmvm().hmvm_hmatrix_write(str(path).encode(), blocks.ctypes.data if len(blocks) else None,
                                        len(blocks), int(nrows), int(ncols),
                                        arena.ctypes.data if arena.size else None, arena.size),
            "hmvm_hmatrix_write")


def load_hmatrix(path) -> StoredHMatrix:
    """Read an H-matrix file (hmvm_hmatrix_read_info + hmvm_hmatrix_read)."""
    nr, nc, nb, na = (C.c_int64() for _ in range(4))
    p = str(path).encode()
    N.check(N.hmvm().hmvm_hmatrix_read_info(p, C.byref(nr), C.byref(nc), C.byref(nb), C.byref(na)),
            "hmvm_hmatrix_read_info")
    blocks = np.zeros(nb.value, dtype=N.block_dtype())
    arena = np.zeros(na.value)
    N.check(N.hmvm().hmvm_hmatrix_read(p, blocks.ctypes.data if nb.value else None, nb.value,
                                       arena.ctypes.data if na.value else None, na.value), "hmvm_hmatrix_read")
    return StoredHMatrix(blocks, arena, int(nr.value), int(nc.value))
