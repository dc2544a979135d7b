"""ctypes bindings of include/hmvm.h and include/hlrgen.h.

The libraries are loaded from hlr_amd/lib/ (built in-tree by hlr_amd/build.py).
A missing library is an error: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIBDIR = Path(__file__).resolve().parent / "lib"

# ---------------------------------------------------------------------------
# constants (hmvm.h)
# ---------------------------------------------------------------------------
HMVM_OK = 0
ERRORS = {
    1: "HMVM_ERR_INVALID_ARG", 2: "HMVM_ERR_BAD_BLOCK", 3: "HMVM_ERR_OVERLAP",
    4: "HMVM_ERR_NONFINITE", 5: "HMVM_ERR_ALLOC", 6: "HMVM_ERR_DEVICE",
    7: "HMVM_ERR_UNSUPPORTED", 8: "HMVM_ERR_LAYOUT", 9: "HMVM_ERR_KERNEL",
    10: "HMVM_ERR_IO", 11: "HMVM_ERR_CORRUPT", 12: "HMVM_ERR_TOO_LARGE",
}
OP_NORMAL, OP_TRANSPOSED, OP_ADJOINT, OP_CONJUGATE = 0, 1, 2, 3
BLOCK_DENSE, BLOCK_LOWRANK, BLOCK_LOWRANK_SV = 0, 1, 2
PREC_FP64, PREC_FP32, PREC_VALR, PREC_BFP, PREC_AFL = 0, 1, 2, 3, 4
F64, F32 = 0, 1
# size of the fixed device scratch of hmvm_apply_noalloc (HMVM_NOALLOC_MAX_BYTES)
NOALLOC_MAX_BYTES = 64 * 1024 * 1024
FMT_F64, FMT_F32, FMT_BF16, FMT_FP8 = 0, 1, 2, 3
FMT_B8, FMT_B16, FMT_B25 = 4, 5, 6
FMT_A12, FMT_A16, FMT_A24 = 7, 8, 9
NFMT = 10
# bits per stored value of each format; a 16-byte row holds 128 // bits values
FMT_BITS = (64, 32, 16, 8, 8, 16, 25, 12, 16, 24)

KERNEL_LOG, KERNEL_LOG_GALERKIN, KERNEL_MATERN, KERNEL_INVR = 0, 1, 2, 3
GEOM_LINE, GEOM_GRID1D, GEOM_CUBE, GEOM_SPHERE = 0, 1, 2, 3
FORMAT_H, FORMAT_TLR = 0, 1
LR_UV, LR_USV = 0, 1


class HmvmBlock(C.Structure):
    _fields_ = [("row_ofs", C.c_int64), ("col_ofs", C.c_int64), ("nrows", C.c_int32),
                ("ncols", C.c_int32), ("rank", C.c_int32), ("kind", C.c_int32),
                ("off_a", C.c_int64), ("off_b", C.c_int64), ("off_s", C.c_int64)]


class HmvmPlanOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("dense_prec", C.c_int32), ("eps", C.c_double),
                ("check_finite", C.c_int32), ("force_format", C.c_int32)]


class HmvmPlanInfo(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nblocks", C.c_int64),
                ("ndense", C.c_int64), ("nlowrank", C.c_int64), ("nrow_panels", C.c_int64),
                ("ncol_panels", C.c_int64), ("factor_bytes", C.c_int64),
                ("bytes_by_format", C.c_int64 * NFMT), ("cols_by_format", C.c_int64 * NFMT),
                ("scale_bytes", C.c_int64), ("flops", C.c_double), ("datasize", C.c_double),
                ("index_bytes", C.c_int64), ("one_level", C.c_int32), ("reserved", C.c_int32),
                ("bytes_rpass", C.c_double), ("bytes_vpass", C.c_double)]


class HlrgenConfig(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("geometry", C.c_int32), ("n", C.c_int64),
                ("seed", C.c_uint64), ("ntile", C.c_int32), ("format", C.c_int32),
                ("eta", C.c_double), ("eps", C.c_double), ("matern_ell", C.c_double),
                ("matern_sigma", C.c_double), ("nn_diag", C.c_int32), ("lrformat", C.c_int32),
                ("nthreads", C.c_int32), ("tlr_std_adm", C.c_int32), ("verbose", C.c_int32),
                ("skip_lowrank", C.c_int32)]


# numpy dtype matching hmvm_block
def block_dtype():
    import numpy as np
    return np.dtype([("row_ofs", "<i8"), ("col_ofs", "<i8"), ("nrows", "<i4"), ("ncols", "<i4"),
                     ("rank", "<i4"), ("kind", "<i4"), ("off_a", "<i8"), ("off_b", "<i8"),
                     ("off_s", "<i8")], align=True)


# numpy dtypes matching hmvm_cluster and hmvm_ublock
BLOCK_COUPLING = 3


def cluster_dtype():
    import numpy as np
    return np.dtype([("ofs", "<i8"), ("size", "<i4"), ("rank", "<i4"), ("parent", "<i4"), ("pad_", "<i4"),
                     ("off_basis", "<i8"), ("off_transfer", "<i8"), ("off_s", "<i8")], align=True)


def ublock_dtype():
    import numpy as np
    return np.dtype([("row_ofs", "<i8"), ("col_ofs", "<i8"), ("nrows", "<i4"), ("ncols", "<i4"),
                     ("kind", "<i4"), ("rowcb", "<i4"), ("colcb", "<i4"), ("pad_", "<i4"), ("off", "<i8")],
                    align=True)


class HmvmUplanInfo(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nrow_clusters", C.c_int64),
                ("ncol_clusters", C.c_int64), ("ncoupling", C.c_int64), ("ndense", C.c_int64),
                ("nested", C.c_int32), ("levels", C.c_int32), ("basis_bytes", C.c_int64),
                ("transfer_bytes", C.c_int64), ("coupling_bytes", C.c_int64), ("dense_bytes", C.c_int64),
                ("flops", C.c_double), ("datasize", C.c_double)]


KFN_LOG, KFN_LOG_GALERKIN, KFN_MATERN, KFN_INVR = 0, 1, 2, 3


class HmvmKernelFn(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("n", C.c_int64), ("h", C.c_double),
                ("ell", C.c_double), ("sigma", C.c_double), ("coords", C.c_void_p), ("perm", C.c_void_p)]


# numpy dtype matching hmvm_lr_leaf
def lr_leaf_dtype():
    import numpy as np
    return np.dtype([("row0", "<i8"), ("col0", "<i8"), ("nrows", "<i4"), ("ncols", "<i4")], align=True)


class HmvmError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, code)}" if what else ERRORS.get(code, str(code)))


def check(rc: int, what: str = "") -> None:
    if rc != HMVM_OK:
        raise HmvmError(rc, what)


_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int

# (name, restype, argtypes)
HMVM_SYMBOLS = [
    ("hmvm_plan_create", _i32, [_vp, _i64, _i64, _i64, _vp, _i64, C.POINTER(HmvmPlanOpts), _i32, C.POINTER(_vp)]),
    ("hmvm_plan_destroy", _i32, [_vp]),
    ("hmvm_plan_get_info", _i32, [_vp, C.POINTER(HmvmPlanInfo)]),
    ("hmvm_apply", _i32, [_vp, C.c_double, _i32, _vp, _i64, _vp, _i64, _i32, _i32, _vp]),
    ("hmvm_apply_timed", _i32, [_vp, C.c_double, _i32, _vp, _vp, _i32, _vp, C.POINTER(C.c_float)]),
    ("hmvm_workspace_size", _i32, [_vp, _i32, _i32, C.POINTER(_i64)]),
    ("hmvm_apply_ws", _i32, [_vp, C.c_double, _i32, _vp, _i64, _vp, _i64, _i32, _i32, _vp, _i64, _vp]),
    ("hmvm_apply_noalloc", _i32, [_vp, C.c_double, _i32, _vp, _i64, _vp, _i64, _i32, _i32, _vp]),
    ("hmvm_tlr_apply", _i32, [_vp, C.c_double, _i32, _vp, _vp, _i32, _vp]),
    ("hmvm_grad", _i32, [_vp, _vp, _vp, _vp, C.POINTER(_vp), _i32, _vp]),
    ("hmvm_plan_arena_elems", _i32, [_vp, _i32, C.POINTER(_i64)]),
    ("hmvm_plan_get_arena", _i32, [_vp, _i32, _vp, _vp]),
    ("hmvm_plan_set_arena", _i32, [_vp, _i32, _vp, _vp]),
    ("hmvm_plan_unpack_block", _i32, [_vp, _i64, C.POINTER(_vp), _i32, _vp, _vp]),
    ("hmvm_plan_arena_source", _i32, [_vp, _i32, _vp]),
    ("hmvm_batch_create", _i32, [C.POINTER(_vp), _i32, C.POINTER(_i32), _i32, C.POINTER(_vp)]),
    ("hmvm_batch_destroy", _i32, [_vp]),
    ("hmvm_batch_apply", _i32, [_vp, C.POINTER(C.c_double), C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_vp),
                                C.POINTER(_i64), _vp]),
    ("hmvm_batch_capture", _i32, [_vp, C.POINTER(C.c_double), C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_vp),
                                  C.POINTER(_i64), _vp]),
    ("hmvm_batch_replay", _i32, [_vp, _vp]),
    ("hmvm_batch_check", _i32, [_vp, _vp]),
    ("hmvm_batch_tune", _i32, [_vp, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_vp), C.POINTER(_i64), _vp,
                               C.POINTER(_i32)]),
    ("hmvm_spectral_norm", _i32, [_vp, C.c_double, _i64, _i32, _vp, C.POINTER(C.c_double), C.POINTER(_i64)]),
    ("hmvm_inv_error_2", _i32, [_vp, _vp, C.c_double, _i64, _vp, C.POINTER(C.c_double), C.POINTER(_i64)]),
    ("hmvm_linear_iteration", _i32, [_vp, _vp, _vp, _vp, _vp, C.c_double, _i64, _vp, C.POINTER(_i64),
                                     C.POINTER(C.c_double)]),
    ("hmvm_hmatrix_write", _i32, [C.c_char_p, _vp, _i64, _i64, _i64, _vp, _i64]),
    ("hmvm_hmatrix_read_info", _i32, [C.c_char_p, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64),
                                      C.POINTER(_i64)]),
    ("hmvm_hmatrix_read", _i32, [C.c_char_p, _vp, _i64, _vp, _i64]),
    ("hmvm_plan_create_from_file", _i32, [C.c_char_p, _vp, _i32, C.POINTER(_vp)]),
    ("hmvm_uplan_create", _i32, [_vp, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _vp, _i64, C.POINTER(HmvmPlanOpts),
                                  _i32, C.POINTER(_vp)]),
    ("hmvm_uplan_destroy", _i32, [_vp]),
    ("hmvm_uplan_get_info", _i32, [_vp, C.POINTER(HmvmUplanInfo)]),
    ("hmvm_uplan_apply", _i32, [_vp, C.c_double, _i32, _vp, _vp, _vp]),
    ("hmvm_lowrank_build", _i32, [C.POINTER(HmvmKernelFn), _vp, _i64, C.c_double, _i32, _i32, C.POINTER(_vp)]),
    ("hmvm_lowrank_ranks", _i32, [_vp, _vp]),
    ("hmvm_lowrank_copy", _i32, [_vp, _vp, _vp, _vp, _i32, _vp, _i64]),
    ("hmvm_lowrank_seconds", C.c_double, [_vp]),
    ("hmvm_lowrank_timing", _i32, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("hmvm_lowrank_free", None, [_vp]),
    ("hmvm_error_string", C.c_char_p, [_i32]),
]

HLRGEN_SYMBOLS = [
    ("hlrgen_default_config", None, [C.POINTER(HlrgenConfig)]),
    ("hlrgen_build", _i32, [C.POINTER(HlrgenConfig), C.POINTER(_vp)]),
    ("hlrgen_free", None, [_vp]),
    ("hlrgen_n", _i64, [_vp]),
    ("hlrgen_nblocks", _i64, [_vp]),
    ("hlrgen_blocks", _vp, [_vp]),
    ("hlrgen_arena_len", _i64, [_vp]),
    ("hlrgen_arena", _vp, [_vp]),
    ("hlrgen_perm", _vp, [_vp]),
    ("hlrgen_coords", _vp, [_vp]),
    ("hlrgen_build_seconds", C.c_double, [_vp]),
    ("hlrgen_eval", _i32, [_vp, _i64, _vp, _i64, _vp, _vp]),
    ("hlrgen_random_vector", None, [C.c_uint64, _i64, C.c_double, C.c_double, _i32, _vp]),
    ("hlrgen_qr", _i32, [_i64, _i64, _vp, _vp, _vp]),
    ("hlrgen_svd", _i32, [_i64, _vp, _vp, _vp, _vp]),
    ("hlrgen_truncate", _i32, [_i64, _i64, _i64, _vp, _vp, C.c_double, _vp, _vp, _vp, C.POINTER(_i64)]),
    ("hlrgen_aca", _i32, [_vp, _i64, _vp, _i64, _vp, C.c_double, _i64, _vp, _vp, C.POINTER(_i64)]),
]


def _load(name: str, symbols) -> C.CDLL:
    path = LIBDIR / name
    if not path.exists():
        raise ImportError(f"{path} is missing: run __graft_entry__.build() (or python hlr_amd/build.py)")
    lib = C.CDLL(str(path))
    for sym, res, args in symbols:
        fn = getattr(lib, sym)
        fn.restype = res
        fn.argtypes = args
    return lib


_hmvm = None
_hlrgen = None


def hmvm() -> C.CDLL:
    global _hmvm
    if _hmvm is None:
        # torch ships its own HIP runtime under the same soname: load it first,
        # so that the library binds to it instead of loading a second runtime
        # (two runtimes in one process see no devices in one of them)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _hmvm = _load("libhmvm.so", HMVM_SYMBOLS)
    return _hmvm


def hlrgen() -> C.CDLL:
    global _hlrgen
    if _hlrgen is None:
        _hlrgen = _load("libhlrgen.so", HLRGEN_SYMBOLS)
    return _hlrgen


def error_string(code: int) -> str:
    return hmvm().hmvm_error_string(code).decode()
