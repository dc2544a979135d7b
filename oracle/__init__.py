"""Oracle: CPU FP64 restatement of HLR's mul_vec path (test infrastructure).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this package, and only as the checker or the timed CPU baseline.
Each function cites the reference code it restates in oracle/hmvm_oracle.c.

Parity status of the functions (what pins each one is listed in the C file
header and exercised by tests/test_oracle.py):
  mul_vec, tlr_mul_vec, mul_vec_chunk      pinned (dense expansion, brute force,
                                           invariants)
  mul_vec_flops, mul_vec_datasize          pinned (hand-counted trees)
  valr_tolerances, eps_to_rate             pinned (closed forms on synthetic
                                           singular values)
  round_bf16, round_e4m3                   pinned (torch dtype conversions and
                                           the OCP E4M3 code table)
  bfp_exp, bfp_round, bfp_format           pinned (numpy rint, brute-force
                                           minimality, representable columns)
  afl_scale, afl_encode, afl_decode        pinned (exactly representable grid
                                           values, truncation onto the grid by
                                           brute force, scale rule, zero code)
  afl_format                               pinned (brute-force minimality over
                                           the pinned rounding, fallbacks)
  uniform_mul_vec, uniform_dense_expand    pinned (numpy products of a
                                           hand-assembled expansion; H^2 equal
                                           to the uniform-H matrix with the
                                           expanded bases)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB = None
_PATH = Path(__file__).resolve().parent / "liboracle.so"


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not _PATH.exists():
            from . import build as _b
            _b.build()
        L = C.CDLL(str(_PATH))
        vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int, C.c_double
        L.oracle_mul_vec.argtypes = [vp, i64, vp, d, i32, vp, i64, vp, i64, i32]
        L.oracle_mul_vec.restype = i32
        L.oracle_tlr_mul_vec.argtypes = [vp, i64, vp, d, i32, vp, vp, i32]
        L.oracle_tlr_mul_vec.restype = i32
        L.oracle_mul_vec_chunk.argtypes = [vp, i64, vp, d, i32, vp, vp, i64, i32]
        L.oracle_mul_vec_chunk.restype = i32
        L.oracle_time_chunk.argtypes = [vp, i64, vp, vp, vp, i64, i32, i32]
        L.oracle_time_chunk.restype = d
        L.oracle_mul_vec_flops.argtypes = [vp, i64, i32]
        L.oracle_mul_vec_flops.restype = d
        L.oracle_mul_vec_datasize.argtypes = [vp, i64, i64, i64]
        L.oracle_mul_vec_datasize.restype = d
        L.oracle_dense_expand.argtypes = [vp, i64, vp, i64, i64, vp]
        L.oracle_dense_expand.restype = i32
        L.oracle_valr_tolerances.argtypes = [i64, vp, d, vp]
        L.oracle_valr_tolerances.restype = None
        L.oracle_eps_to_rate.argtypes = [d]
        L.oracle_eps_to_rate.restype = i32
        L.oracle_format_for_rate.argtypes = [i32]
        L.oracle_format_for_rate.restype = i32
        L.oracle_round_bf16.argtypes = [d]
        L.oracle_round_bf16.restype = d
        L.oracle_encode_e4m3.argtypes = [d]
        L.oracle_encode_e4m3.restype = C.c_uint8
        L.oracle_decode_e4m3.argtypes = [C.c_uint8]
        L.oracle_decode_e4m3.restype = d
        L.oracle_round_e4m3.argtypes = [d]
        L.oracle_round_e4m3.restype = d
        L.oracle_bfp_exp.argtypes = [d, i32]
        L.oracle_bfp_exp.restype = i32
        L.oracle_bfp_round.argtypes = [d, i32, i32]
        L.oracle_bfp_round.restype = d
        L.oracle_bfp_format.argtypes = [i32, vp, vp, vp, d, i32]
        L.oracle_bfp_format.restype = i32
        L.oracle_afl_scale.argtypes = [vp, i64]
        L.oracle_afl_scale.restype = C.c_float
        L.oracle_afl_encode.argtypes = [d, C.c_float, i32]
        L.oracle_afl_encode.restype = C.c_uint32
        L.oracle_afl_decode.argtypes = [C.c_uint32, C.c_float, i32]
        L.oracle_afl_decode.restype = d
        L.oracle_afl_round.argtypes = [d, C.c_float, i32]
        L.oracle_afl_round.restype = d
        L.oracle_afl_format.argtypes = [i32, vp, vp, vp, d, i32]
        L.oracle_afl_format.restype = i32
        L.oracle_uniform_mul_vec.argtypes = [vp, i64, vp, i64, vp, i64, vp, d, i32, vp, vp]
        L.oracle_uniform_mul_vec.restype = i32
        L.oracle_uniform_dense_expand.argtypes = [vp, i64, vp, i64, vp, i64, vp, i64, i64, vp]
        L.oracle_uniform_dense_expand.restype = i32
        _LIB = L
    return _LIB


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


def mul_vec(blocks, arena, alpha, op, x, y):
    """y += alpha op(A) x in the sequential leaf order (y modified in place).
    x, y: 1-D, or 2-D (nrhs, n) C-contiguous arrays."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert y.dtype == np.float64 and y.flags.c_contiguous
    nrhs = x.shape[0] if x.ndim == 2 else 1
    ldx = x.shape[-1]
    ldy = y.shape[-1]
    rc = lib().oracle_mul_vec(_p(blocks), len(blocks), _p(arena), float(alpha), int(op), _p(x), ldx, _p(y), ldy, nrhs)
    assert rc == 0, rc
    return y


def tlr_mul_vec(blocks, arena, alpha, op, x, y, nthreads=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    rc = lib().oracle_tlr_mul_vec(_p(blocks), len(blocks), _p(arena), float(alpha), int(op), _p(x), _p(y), nthreads)
    assert rc == 0, rc
    return y


def mul_vec_chunk(blocks, arena, alpha, op, x, y, nthreads=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    rc = lib().oracle_mul_vec_chunk(_p(blocks), len(blocks), _p(arena), float(alpha), int(op), _p(x), _p(y),
                                    len(y), nthreads)
    assert rc == 0, rc
    return y


def time_chunk(blocks, arena, x, y, nthreads, nmvm):
    return lib().oracle_time_chunk(_p(blocks), len(blocks), _p(arena), _p(x), _p(y), len(y), nthreads, nmvm)


def mul_vec_flops(blocks, op=0):
    return lib().oracle_mul_vec_flops(_p(blocks), len(blocks), op)


def mul_vec_datasize(blocks, nrows, ncols):
    return lib().oracle_mul_vec_datasize(_p(blocks), len(blocks), nrows, ncols)


def dense_expand(blocks, arena, nrows, ncols):
    A = np.zeros((ncols, nrows))      # column-major nrows x ncols
    rc = lib().oracle_dense_expand(_p(blocks), len(blocks), _p(arena), nrows, ncols, _p(A))
    assert rc == 0, rc
    return A.T


def valr_tolerances(sv, eps):
    sv = np.ascontiguousarray(sv, dtype=np.float64)
    out = np.zeros_like(sv)
    lib().oracle_valr_tolerances(len(sv), _p(sv), float(eps), _p(out))
    return out


def eps_to_rate(eps):
    return lib().oracle_eps_to_rate(float(eps))


def format_for_rate(rate):
    return lib().oracle_format_for_rate(int(rate))


def round_bf16(v):
    return lib().oracle_round_bf16(float(v))


def encode_e4m3(v):
    return lib().oracle_encode_e4m3(float(v))


def decode_e4m3(c):
    return lib().oracle_decode_e4m3(int(c))


def round_e4m3(v):
    return lib().oracle_round_e4m3(float(v))


def bfp_exp(maxabs, nb):
    return lib().oracle_bfp_exp(float(maxabs), int(nb))


def bfp_round(v, k, nb):
    return lib().oracle_bfp_round(float(v), int(k), int(nb))


def _pieces(n, piece):
    if np.isscalar(piece):
        return [piece] * (n // piece) + ([n % piece] if n % piece else [])
    return list(piece)


def bfp_round_column(col, nb, piece=64):
    """Stored values of a column in panel columns of `piece` rows (or the
    given list of piece lengths), each with its own step."""
    col = np.asarray(col, dtype=np.float64)
    out = np.empty_like(col)
    r0 = 0
    for ln in _pieces(len(col), piece):
        seg = col[r0:r0 + ln]
        k = bfp_exp(np.abs(seg).max() if ln else 0.0, nb)
        out[r0:r0 + ln] = [bfp_round(v, k, nb) for v in seg]
        r0 += ln
    return out


def bfp_format(cols, tol, rate, piece=64):
    """Format chosen for the columns `cols` (list of 1-D arrays) sharing one
    width under relative 2-norm tolerance tol (fallback by rate), each cut
    into panel columns of `piece` rows (or one list of piece lengths per
    column)."""
    vals, npieces, plen = [], [], []
    for c, col in enumerate(cols):
        col = np.asarray(col, dtype=np.float64)
        lens = _pieces(len(col), piece if np.isscalar(piece) else piece[c])
        vals.append(col)
        npieces.append(len(lens))
        plen.extend(lens)
    v = np.ascontiguousarray(np.concatenate(vals) if vals else np.zeros(0))
    npc = np.ascontiguousarray(npieces, dtype=np.int64)
    pl = np.ascontiguousarray(plen, dtype=np.int64)
    return lib().oracle_bfp_format(len(cols), _p(v), _p(npc), _p(pl), float(tol), int(rate))


def uniform_mul_vec(rowcb, colcb, blocks, arena, alpha, op, x, y):
    """y += alpha op(M) x for a uniform-H / H^2 matrix (hmvm_cluster and
    hmvm_ublock record arrays), y modified in place."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert y.dtype == np.float64 and y.flags.c_contiguous
    rc = lib().oracle_uniform_mul_vec(_p(rowcb), len(rowcb), _p(colcb), len(colcb), _p(blocks), len(blocks),
                                      _p(arena), float(alpha), int(op), _p(x), _p(y))
    assert rc == 0, rc
    return y


def uniform_dense_expand(rowcb, colcb, blocks, arena, nrows, ncols):
    A = np.zeros((ncols, nrows))
    rc = lib().oracle_uniform_dense_expand(_p(rowcb), len(rowcb), _p(colcb), len(colcb), _p(blocks), len(blocks),
                                           _p(arena), nrows, ncols, _p(A))
    assert rc == 0, rc
    return A.T


def afl_scale(col):
    col = np.ascontiguousarray(col, dtype=np.float64)
    return float(lib().oracle_afl_scale(_p(col), col.size))


def afl_encode(v, scale, nb):
    return int(lib().oracle_afl_encode(float(v), float(scale), int(nb)))


def afl_decode(code, scale, nb):
    return float(lib().oracle_afl_decode(int(code), float(scale), int(nb)))


def afl_round_column(col, nb, piece=64):
    """Stored values of a column in AFL codes of nb bits, one scale per
    piece of `piece` rows (or the given list of piece lengths)."""
    col = np.asarray(col, dtype=np.float64)
    out = np.empty_like(col)
    r0 = 0
    for ln in _pieces(len(col), piece):
        seg = col[r0:r0 + ln]
        s = afl_scale(seg)
        out[r0:r0 + ln] = [lib().oracle_afl_round(float(v), s, int(nb)) for v in seg]
        r0 += ln
    return out


def afl_format(cols, tol, rate, piece=64):
    """Format chosen for the columns `cols` sharing one AFL width under
    relative 2-norm tolerance tol (fallback by rate); pieces as bfp_format."""
    vals, npieces, plen = [], [], []
    for c, col in enumerate(cols):
        col = np.asarray(col, dtype=np.float64)
        lens = _pieces(len(col), piece if np.isscalar(piece) else piece[c])
        vals.append(col)
        npieces.append(len(lens))
        plen.extend(lens)
    v = np.ascontiguousarray(np.concatenate(vals) if vals else np.zeros(0))
    npc = np.ascontiguousarray(npieces, dtype=np.int64)
    pl = np.ascontiguousarray(plen, dtype=np.int64)
    return lib().oracle_afl_format(len(cols), _p(v), _p(npc), _p(pl), float(tol), int(rate))
