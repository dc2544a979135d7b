"""H-matrix files (CPU): write / read round trips, the documented layout,
and rejection of damaged files.

The layout is pinned independently of the library: the header fields, the
record words and the three checksums are recomputed here from the format's
description (a reader written from include/hmvm.h alone), and a plan built
from a file equals the plan built from the same leaves in memory."""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np
import pytest

from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, LOG1D, MATERN_TLR, load_hmatrix, save_hmatrix
from hlr_amd.plan import Plan

MASK = (1 << 64) - 1


def fold(words) -> int:
    """The format's checksum over 8-byte words."""
    h = 0x484D564D
    for w in words:
        h = (((h ^ int(w)) * 0x9E3779B97F4A7C15) & MASK) ^ (h >> 29)
    return h


def fold_np(arr: np.ndarray) -> int:
    return fold(np.ascontiguousarray(arr).view(np.uint64).tolist())


RECORD_FIELDS = ("row_ofs", "col_ofs", "nrows", "ncols", "rank", "kind", "off_a", "off_b", "off_s")


def record_words(blocks) -> np.ndarray:
    return np.stack([blocks[f].astype(np.int64) for f in RECORD_FIELDS], axis=1).reshape(-1)


@pytest.fixture(scope="module")
def logm():
    return HMatrix(LOG1D.with_(n=2048))


def test_roundtrip_and_layout(logm, tmp_path):
    path = tmp_path / "log.hm"
    logm.save(path)
    raw = path.read_bytes()
    nb, na = len(logm.blocks), logm.arena.size
    assert len(raw) == 72 + 72 * nb + 8 * na
    assert raw[:8] == b"HMVMHMAT"
    version, zero, nr, nc, nbf, naf, sb, sa, sh = struct.unpack("<II4q3Q", raw[8:72])
    assert (version, zero, nr, nc, nbf, naf) == (1, 0, logm.n, logm.n, nb, na)
    rec = np.frombuffer(raw[72:72 + 72 * nb], dtype="<i8")
    assert np.array_equal(rec, record_words(logm.blocks))
    assert np.array_equal(np.frombuffer(raw[72 + 72 * nb:], dtype="<f8"), logm.arena)
    assert sb == fold(rec.view(np.uint64).tolist())
    assert sa == fold_np(logm.arena)
    assert sh == fold(np.frombuffer(raw[:64], dtype="<u8").tolist())
    back = load_hmatrix(path)
    assert (back.nrows, back.ncols, back.n) == (logm.n, logm.n, logm.n)
    assert back.blocks.tobytes() == np.ascontiguousarray(logm.blocks).tobytes()
    assert back.arena.tobytes() == logm.arena.tobytes()
    # writing what was read gives the same file
    back.save(tmp_path / "again.hm")
    assert (tmp_path / "again.hm").read_bytes() == raw


def test_empty_matrix(tmp_path):
    path = tmp_path / "empty.hm"
    save_hmatrix(path, np.zeros(0, dtype=N.block_dtype()), np.zeros(0), 8, 8)
    assert path.stat().st_size == 72
    m = load_hmatrix(path)
    assert len(m.blocks) == 0 and m.arena.size == 0 and (m.nrows, m.ncols) == (8, 8)
    Plan.from_file(path, device=-1)


def _rewrite_checksums(raw: bytearray, nb: int) -> None:
    rec = np.frombuffer(bytes(raw[72:72 + 72 * nb]), dtype="<u8")
    arena = np.frombuffer(bytes(raw[72 + 72 * nb:]), dtype="<u8")
    raw[48:56] = struct.pack("<Q", fold(rec.tolist()))
    raw[56:64] = struct.pack("<Q", fold(arena.tolist()))
    raw[64:72] = struct.pack("<Q", fold(np.frombuffer(bytes(raw[:64]), dtype="<u8").tolist()))


def _read_rc(path, nb, na):
    blocks = np.zeros(nb, dtype=N.block_dtype())
    arena = np.zeros(na)
    return N.hmvm().hmvm_hmatrix_read(str(path).encode(), blocks.ctypes.data if nb else None, nb,
                                      arena.ctypes.data if na else None, na)


def test_damaged_files_rejected(logm, tmp_path):
    good = tmp_path / "good.hm"
    logm.save(good)
    raw = good.read_bytes()
    nb, na = len(logm.blocks), logm.arena.size
    bad = tmp_path / "bad.hm"

    def case(data: bytes, code: int, sizes=(nb, na)):
        bad.write_bytes(data)
        assert _read_rc(bad, *sizes) == code

    case(raw, N.HMVM_OK)
    flip = lambda i: raw[:i] + bytes([raw[i] ^ 0x10]) + raw[i + 1:]
    case(flip(72 + 72 * nb + 8 * (na // 2) + 3), 11)    # arena byte
    case(flip(72 + 72 * (nb // 2) + 5), 11)             # record byte
    case(flip(20), 11)                                  # header field
    case(flip(60), 11)                                  # stored arena checksum
    case(raw[:-1], 10)                                  # truncated arena
    case(raw[:40], 10)                                  # truncated header
    case(raw + b"\0", 11)                               # data after the arena
    case(b"HMVMHMAX" + raw[8:], 10)                     # not this format
    assert _read_rc(tmp_path / "missing.hm", nb, na) == 10
    case(raw, 1, sizes=(nb - 1, na))                    # caller sizes differ
    case(raw, 1, sizes=(nb, na + 1))
    # an unknown version with consistent checksums
    v2 = bytearray(raw)
    v2[8:12] = struct.pack("<I", 2)
    _rewrite_checksums(v2, nb)
    case(bytes(v2), 10)
    # a record field outside int32 with consistent checksums
    big = bytearray(raw)
    big[72 + 16:72 + 24] = struct.pack("<q", 1 << 40)   # nrows of record 0
    _rewrite_checksums(big, nb)
    case(bytes(big), 11)
    # read_info sees only the header
    nr = C.c_int64()
    assert N.hmvm().hmvm_hmatrix_read_info(str(good).encode(), C.byref(nr), None, None, None) == 0
    assert nr.value == logm.n
    # writing into a missing directory fails and leaves nothing behind
    blocks = np.ascontiguousarray(logm.blocks)
    rc = N.hmvm().hmvm_hmatrix_write(str(tmp_path / "no" / "x.hm").encode(), blocks.ctypes.data, nb, logm.n,
                                     logm.n, logm.arena.ctypes.data, na)
    assert rc == 10 and not (tmp_path / "no").exists()
    assert N.hmvm().hmvm_hmatrix_write(None, None, 0, 1, 1, None, 0) == 1
    assert N.hmvm().hmvm_hmatrix_write(str(bad).encode(), None, 3, 1, 1, None, 0) == 1
    assert N.error_string(10) == "file missing, short, or not an H-matrix file"


@pytest.mark.parametrize("prec", ["fp64", "valr", "bfp"])
def test_plan_from_file_equals_plan_from_memory(tmp_path, prec):
    """hmvm_plan_create_from_file builds the same plan as hmvm_plan_create on
    the leaves in memory: same statistics, bitwise the same arenas."""
    hm = HMatrix(MATERN_TLR.with_(n=4096, lrformat=N.LR_USV))
    path = tmp_path / "tlr.hm"
    hm.save(path)
    ref = Plan.from_hmatrix(hm, precision=prec, eps=1e-6, device=-1)
    opts = N.HmvmPlanOpts()
    opts.precision = {"fp64": N.PREC_FP64, "valr": N.PREC_VALR, "bfp": N.PREC_BFP}[prec]
    opts.dense_prec = N.PREC_VALR
    opts.eps = 1e-6
    opts.check_finite = 1
    p = C.c_void_p()
    N.check(N.hmvm().hmvm_plan_create_from_file(str(path).encode(), C.byref(opts), -1, C.byref(p)))
    try:
        got = N.HmvmPlanInfo()
        exp = N.HmvmPlanInfo()
        N.check(N.hmvm().hmvm_plan_get_info(p, C.byref(got)))
        N.check(N.hmvm().hmvm_plan_get_info(ref.handle, C.byref(exp)))
        assert bytes(got) == bytes(exp)
        for f in range(N.NFMT):
            n = ref.arena_bytes(f)
            if not n:
                continue
            buf = np.zeros(n, dtype=np.uint8)
            N.check(N.hmvm().hmvm_plan_get_arena(p, f, C.c_void_p(buf.ctypes.data), None))
            assert buf.tobytes() == ref.arena_numpy(f).view(np.uint8).tobytes(), f
    finally:
        N.hmvm().hmvm_plan_destroy(p)
    # the Python spelling reads the same matrix
    q = Plan.from_file(path, precision=prec, eps=1e-6, device=-1)
    assert q.info() == ref.info()


def test_offsets_near_int64_max_rejected(logm, tmp_path):
    """A record whose offset would wrap int64 when added to its length is
    rejected by the plan build (HMVM_ERR_BAD_BLOCK), also when the file's
    checksums are recomputed so that the reader accepts it."""
    good = tmp_path / "good.hm"
    logm.save(good)
    raw = bytearray(good.read_bytes())
    nb = len(logm.blocks)
    lr = int(np.nonzero(logm.blocks["kind"] != N.BLOCK_DENSE)[0][0])
    dn = int(np.nonzero(logm.blocks["kind"] == N.BLOCK_DENSE)[0][0])
    for b, field in ((lr, "off_a"), (lr, "off_b"), (dn, "off_a"), (lr, "row_ofs"), (lr, "col_ofs")):
        bad = bytearray(raw)
        o = 72 + 72 * b + 8 * RECORD_FIELDS.index(field)
        bad[o:o + 8] = struct.pack("<q", (1 << 63) - 5)
        _rewrite_checksums(bad, nb)
        path = tmp_path / f"wrap_{field}.hm"
        path.write_bytes(bytes(bad))
        assert _read_rc(path, nb, logm.arena.size) == N.HMVM_OK      # the file itself is consistent
        p = C.c_void_p()
        opts = N.HmvmPlanOpts()
        opts.precision, opts.check_finite = N.PREC_FP64, 1
        rc = N.hmvm().hmvm_plan_create_from_file(str(path).encode(), C.byref(opts), -1, C.byref(p))
        assert rc == 2, (field, rc)                                   # HMVM_ERR_BAD_BLOCK
