"""Host-side checks that need no GPU: the C-ABI libraries load and export
every declared symbol, plan layout and storage precision on host-only plans
(device = -1) match the oracle bit for bit, and invalid input is rejected
with the documented error codes."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.hmatrix import HMatrix, INVR_SPHERE, LOG1D, MATERN_TLR
from hlr_amd.plan import Plan

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:hmvm|hlrgen)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("hmvm.h", "libhmvm.so"), ("hlrgen.h", "libhlrgen.so")])
def test_exports(header, lib):
    so = C.CDLL(str(ROOT / "hlr_amd" / "lib" / lib))
    names = _declared(header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    names = _declared("hmvm.h")
    bound = {s[0] for s in N.HMVM_SYMBOLS}
    assert set(names) <= bound, set(names) - bound


@pytest.fixture(scope="module")
def sphere():
    return HMatrix(INVR_SPHERE.with_(n=4096))


def test_fp64_layout_roundtrip(sphere):
    p = Plan.from_hmatrix(sphere, precision="fp64", device=-1)
    ar = [p.arena_numpy(f) for f in range(N.NFMT)]
    for b, blk in enumerate(sphere.blocks):
        m, n, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
        A, B = p.unpack_block(b, ar)
        if int(blk["kind"]) == N.BLOCK_DENSE:
            assert np.array_equal(A, sphere.arena[blk["off_a"]:blk["off_a"] + m * n].reshape(n, m).T)
        elif k:
            assert np.array_equal(A, sphere.arena[blk["off_a"]:blk["off_a"] + m * k].reshape(k, m).T)
            assert np.array_equal(B, sphere.arena[blk["off_b"]:blk["off_b"] + n * k].reshape(k, n).T)
    info = p.info()
    assert info["datasize"] == pytest.approx(oracle.mul_vec_datasize(sphere.blocks, sphere.n, sphere.n))
    assert info["flops"] == pytest.approx(oracle.mul_vec_flops(sphere.blocks))


@pytest.mark.parametrize("eps", [1e-4, 1e-6, 1e-8])
def test_valr_chooser_and_quantisation(sphere, eps):
    """Per-column formats follow the oracle's VALR chooser; every stored
    value equals the oracle's rounding of the FP64 value (after the
    power-of-two scale for FP8), bit for bit."""
    p = Plan.from_hmatrix(sphere, precision="valr", eps=eps, device=-1)
    ar = [p.arena_numpy(f) for f in range(N.NFMT)]
    counts = [0] * N.NFMT
    rng = np.random.default_rng(3)
    for b, blk in enumerate(sphere.blocks):
        m, n, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
        if int(blk["kind"]) == N.BLOCK_DENSE:
            continue
        S = sphere.arena[blk["off_s"]:blk["off_s"] + k]
        fm = [oracle.format_for_rate(oracle.eps_to_rate(u)) for u in oracle.valr_tolerances(S, eps)]
        for f in fm:
            counts[f] += 1
        if rng.random() > 0.1:
            continue
        U0 = sphere.arena[blk["off_a"]:blk["off_a"] + m * k].reshape(k, m).T
        V0 = sphere.arena[blk["off_b"]:blk["off_b"] + n * k].reshape(k, n).T
        U, V = p.unpack_block(b, ar)
        for l, f in enumerate(fm):
            for got, ref, rows in ((U[:, l], U0[:, l], m), (V[:, l], V0[:, l], n)):
                if f == 1:
                    exp = ref.astype(np.float32).astype(np.float64)
                elif f == 2:
                    exp = np.array([oracle.round_bf16(v) for v in ref])
                elif f == 3:
                    # scale per 64-row panel column: largest power of two step with max|v|/s <= 448
                    exp = np.empty(rows)
                    for r0 in range(0, rows, 64):
                        seg = ref[r0:r0 + 64]
                        mx = np.abs(seg).max()
                        s = 2.0 ** np.ceil(np.log2(mx / 448.0)) if mx > 0 else 1.0
                        if mx / s > 448:
                            s *= 2
                        exp[r0:r0 + 64] = [oracle.round_e4m3(float(np.float32(v / s))) * s for v in seg]
                else:
                    exp = ref
                assert np.array_equal(got, exp), (b, l, f)
    assert counts == p.info()["cols_by_format"]


def _packed_expected(ref, f):
    if N.FMT_B8 <= f <= N.FMT_B25:
        return oracle.bfp_round_column(ref, N.FMT_BITS[f])
    if f >= N.FMT_A12:
        return oracle.afl_round_column(ref, N.FMT_BITS[f])
    return ref.astype(np.float32).astype(np.float64) if f == N.FMT_F32 else ref


@pytest.mark.parametrize("prec", ["bfp", "afl"])
@pytest.mark.parametrize("eps", [1e-4, 1e-6, 1e-8])
def test_packed_chooser_and_quantisation(sphere, prec, eps):
    """Block floating point and AFL plans: every column's width follows the
    oracle's chooser (U and V of a low-rank column share one width within
    the column's VALR tolerance, dense columns within 2^-(rate(eps)+1)), and
    every stored value equals the oracle's rounding with one step (BFP) or
    scale (AFL) per 64-row panel column, bit for bit."""
    choose = oracle.bfp_format if prec == "bfp" else oracle.afl_format
    first, last = (N.FMT_B8, N.FMT_B25) if prec == "bfp" else (N.FMT_A12, N.FMT_A24)
    p = Plan.from_hmatrix(sphere, precision=prec, eps=eps, device=-1)
    ar = [p.arena_numpy(f) for f in range(N.NFMT)]
    counts = [0] * N.NFMT
    rng = np.random.default_rng(5)
    rate_eps = oracle.eps_to_rate(eps)
    for b, blk in enumerate(sphere.blocks):
        m, n, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
        check = rng.random() < 0.1
        if int(blk["kind"]) == N.BLOCK_DENSE:
            if not check:
                continue
            D0 = sphere.arena[blk["off_a"]:blk["off_a"] + m * n].reshape(n, m).T
            D, _ = p.unpack_block(b, ar)
            for j in range(n):
                f = choose([D0[:, j]], 2.0 ** -(rate_eps + 1), rate_eps)
                assert np.array_equal(D[:, j], _packed_expected(D0[:, j], f)), (b, j, f)
            continue
        U0 = sphere.arena[blk["off_a"]:blk["off_a"] + m * k].reshape(k, m).T
        V0 = sphere.arena[blk["off_b"]:blk["off_b"] + n * k].reshape(k, n).T
        S = sphere.arena[blk["off_s"]:blk["off_s"] + k]
        fm = [choose([U0[:, l], V0[:, l]], u, oracle.eps_to_rate(u))
              for l, u in enumerate(oracle.valr_tolerances(S, eps))]
        for f in fm:
            counts[f] += 1
        if not check:
            continue
        U, V = p.unpack_block(b, ar)
        for l, f in enumerate(fm):
            assert np.array_equal(U[:, l], _packed_expected(U0[:, l], f)), (b, l, f)
            assert np.array_equal(V[:, l], _packed_expected(V0[:, l], f)), (b, l, f)
    info = p.info()
    assert counts == info["cols_by_format"]
    # the packed widths carry the columns (AFL's widest code has 18 mantissa
    # bits, short of the tightest columns at eps = 1e-8), and the plan is
    # smaller than the mixed hardware-format plan at the same eps
    share = 0.5 if (prec == "afl" and eps < 1e-7) else 0.8
    assert sum(info["cols_by_format"][first:last + 1]) > share * sum(counts)
    q = Plan.from_hmatrix(sphere, precision="valr", eps=eps, device=-1)
    assert info["factor_bytes"] < q.info()["factor_bytes"]
    # bytes of each arena: 16 per row of 128 // bits values
    for f in range(N.NFMT):
        assert info["bytes_by_format"][f] == p.arena_elems(f) // (128 // N.FMT_BITS[f]) * 16


def test_packed_force_format_and_errors(sphere):
    """force_format stores every low-rank column in one width; bad options
    are rejected."""
    names = ("bfp8", "bfp16", "bfp25", "afl12", "afl16", "afl24")
    for f in range(N.FMT_B8, N.NFMT):
        prec = "bfp" if f <= N.FMT_B25 else "afl"
        p = Plan.from_hmatrix(sphere, precision=prec, eps=1e-6, device=-1, force_format=names[f - N.FMT_B8])
        c = p.info()["cols_by_format"]
        assert c[f] == sum(c) > 0
    for prec in ("bfp", "afl"):
        with pytest.raises(RuntimeError):
            Plan.from_hmatrix(sphere, precision=prec, eps=0.0, device=-1)
        # BF16 / FP8 columns may not sit beside packed dense leaves
        with pytest.raises(RuntimeError):
            Plan.from_hmatrix(sphere, precision=prec, eps=1e-6, device=-1, force_format="bf16")
    # nor block floating point beside AFL
    with pytest.raises(RuntimeError):
        Plan.from_hmatrix(sphere, precision="afl", eps=1e-6, device=-1, force_format="bfp16")


def test_dense_and_datasize_valr(sphere):
    p = Plan.from_hmatrix(sphere, precision="valr", eps=1e-6, device=-1)
    ar = [p.arena_numpy(f) for f in range(N.NFMT)]
    for b, blk in enumerate(sphere.blocks[:200]):
        if int(blk["kind"]) == N.BLOCK_DENSE:
            m, n = int(blk["nrows"]), int(blk["ncols"])
            D, _ = p.unpack_block(b, ar)
            ref = sphere.arena[blk["off_a"]:blk["off_a"] + m * n].reshape(n, m).T
            assert np.array_equal(D, ref.astype(np.float32).astype(np.float64))


def test_tlr_one_level():
    hm = HMatrix(MATERN_TLR.with_(n=2048, ntile=256))
    assert Plan.from_hmatrix(hm, device=-1).info()["one_level"] == 1
    hm2 = HMatrix(LOG1D.with_(n=1024))
    assert Plan.from_hmatrix(hm2, device=-1).info()["one_level"] == 0


# --- error handling ----------------------------------------------------------

def _one_block(**kw):
    bt = N.block_dtype()
    b = np.zeros(1, dtype=bt)
    b["row_ofs"], b["col_ofs"], b["nrows"], b["ncols"] = 0, 0, 4, 4
    b["kind"], b["rank"], b["off_a"], b["off_b"], b["off_s"] = N.BLOCK_LOWRANK, 2, 0, 8, -1
    for k, v in kw.items():
        b[k] = v
    return b


@pytest.mark.parametrize("kw,code", [
    ({"rank": 5}, 2),                      # rank > min(m, n)
    ({"row_ofs": 2}, 2),                   # block leaves the matrix
    ({"off_b": 100}, 2),                   # factor outside the arena
    ({"kind": 7}, 2),                      # unknown leaf kind
    ({"nrows": -1}, 2),
])
def test_bad_blocks(kw, code):
    arena = np.ones(16)
    with pytest.raises(N.HmvmError) as e:
        Plan(_one_block(**kw), arena, 4, 4, device=-1)
    assert e.value.code == code


def test_nonfinite_and_overlap():
    arena = np.ones(16)
    arena[3] = np.nan
    with pytest.raises(N.HmvmError) as e:
        Plan(_one_block(), arena, 4, 4, device=-1)
    assert e.value.code == 4
    Plan(_one_block(), arena, 4, 4, device=-1, check_finite=False)      # allowed when unchecked
    two = np.concatenate([_one_block(), _one_block()])
    with pytest.raises(N.HmvmError) as e:
        Plan(two, np.ones(16), 4, 4, device=-1)
    assert e.value.code == 3


def test_empty_and_rank0():
    # no leaves at all, and a rank-0 leaf: valid plans with nothing to do
    Plan(np.zeros(0, dtype=N.block_dtype()), np.zeros(0), 8, 8, device=-1)
    p = Plan(_one_block(rank=0), np.zeros(16), 4, 4, device=-1)
    assert p.info()["nlowrank"] == 1


def test_apply_on_host_plan_rejected():
    p = Plan(_one_block(), np.ones(16), 4, 4, device=-1)
    rc = N.hmvm().hmvm_apply(p.handle, 1.0, 0, C.c_void_p(8), 0, C.c_void_p(8), 0, 1, N.F64, None)
    assert rc == 1
    rc = N.hmvm().hmvm_apply(None, 1.0, 0, None, 0, None, 0, 1, N.F64, None)
    assert rc == 1
    assert N.error_string(3) == "leaf blocks overlap"


def test_consumers_reject_bad_arguments():
    """The consumers of the product (spectral norm, inversion error, linear
    iteration) return HMVM_ERR_INVALID_ARG for missing plans or vectors,
    negative limits or tolerances, and host-only plans, before any device
    work."""
    lib = N.hmvm()
    p = Plan(_one_block(), np.ones(16), 4, 4, device=-1)
    v = np.zeros(4)
    vp = C.c_void_p(v.ctypes.data)
    out = C.c_double()
    steps = C.c_int64(-7)
    assert lib.hmvm_spectral_norm(p.handle, 0.0, 0, 1, None, C.byref(out), None) == 1
    assert lib.hmvm_spectral_norm(p.handle, 0.0, -1, 1, None, C.byref(out), None) == 1
    assert lib.hmvm_spectral_norm(None, 0.0, 0, 1, None, C.byref(out), None) == 1
    assert lib.hmvm_inv_error_2(p.handle, p.handle, 0.0, 0, None, C.byref(out), None) == 1
    assert lib.hmvm_inv_error_2(p.handle, None, 0.0, 0, None, C.byref(out), None) == 1
    assert lib.hmvm_inv_error_2(p.handle, p.handle, 0.0, 0, None, None, None) == 1
    for args in [(p.handle, p.handle, vp, vp, None, 0.0, 0), (None, p.handle, vp, vp, None, 0.0, 0),
                 (p.handle, p.handle, None, vp, None, 0.0, 0), (p.handle, p.handle, vp, None, None, 0.0, 0),
                 (p.handle, p.handle, vp, vp, None, -1.0, 0), (p.handle, p.handle, vp, vp, None, 0.0, -1)]:
        assert lib.hmvm_linear_iteration(*args, None, C.byref(steps), None) == 1, args
    assert steps.value == -7


@pytest.mark.parametrize("eps,exact", [(1e-4, False), (1e-6, False), (1e-8, True)])
def test_dense_precision_follows_eps(sphere, eps, exact):
    """Dense leaves under the default rule: FP32 while ceil(-log2 eps) <= 23,
    FP64 below (FP32 rounding alone would exceed eps)."""
    p = Plan.from_hmatrix(sphere, precision="valr", eps=eps, device=-1)
    ar = [p.arena_numpy(f) for f in range(N.NFMT)]
    assert (oracle.eps_to_rate(eps) > 23) == exact
    for b, blk in enumerate(sphere.blocks[:300]):
        if int(blk["kind"]) != N.BLOCK_DENSE:
            continue
        m, n = int(blk["nrows"]), int(blk["ncols"])
        D, _ = p.unpack_block(b, ar)
        ref = sphere.arena[blk["off_a"]:blk["off_a"] + m * n].reshape(n, m).T
        assert np.array_equal(D, ref if exact else ref.astype(np.float32).astype(np.float64))
    # an explicit request overrides the rule
    q = Plan.from_hmatrix(sphere, precision="valr", eps=eps, dense_prec="fp32", device=-1)
    qa = [q.arena_numpy(f) for f in range(N.NFMT)]
    b = next(i for i, blk in enumerate(sphere.blocks) if int(blk["kind"]) == N.BLOCK_DENSE)
    D, _ = q.unpack_block(b, qa)
    blk = sphere.blocks[b]
    m, n = int(blk["nrows"]), int(blk["ncols"])
    ref = sphere.arena[blk["off_a"]:blk["off_a"] + m * n].reshape(n, m).T
    assert np.array_equal(D, ref.astype(np.float32).astype(np.float64))


def test_consumer_python_wrappers_check_vectors():
    """linear_iteration takes contiguous float64 vectors of the plan's size
    (checked before the C call); inv_error_2 on host-only plans reports
    HMVM_ERR_INVALID_ARG as an HmvmError."""
    import torch
    from hlr_amd.plan import inv_error_2, linear_iteration
    p = Plan(_one_block(), np.ones(16), 4, 4, device=-1)
    good = torch.zeros(4, dtype=torch.float64)
    for bad in (torch.zeros(4, dtype=torch.float32), torch.zeros(5, dtype=torch.float64),
                torch.zeros(8, dtype=torch.float64)[::2], torch.zeros((2, 2), dtype=torch.float64)):
        with pytest.raises(ValueError):
            linear_iteration(p, p, bad, good)
        with pytest.raises(ValueError):
            linear_iteration(p, p, good, bad)
    with pytest.raises(ValueError):
        linear_iteration(p, p, good, good, sol=torch.zeros(3, dtype=torch.float64))
    with pytest.raises(N.HmvmError) as ei:
        inv_error_2(p, p, stream=_NoStream())
    assert ei.value.code == 1


class _NoStream:
    """stands in for a torch stream on a machine without a GPU"""
    cuda_stream = 0
