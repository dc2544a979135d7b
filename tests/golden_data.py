"""Loader of the exact-arithmetic fixtures in tests/golden/exact_trees.npz
(written by tests/golden/make_golden.py)."""
from __future__ import annotations

import os

import numpy as np

from hlr_amd import _native as N

CASES = ("h_ragged", "tlr")
_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "exact_trees.npz")


def load(case: str) -> dict:
    with np.load(_PATH, allow_pickle=False) as d:
        out = {k[len(case) + 1:]: d[k] for k in d.files if k.startswith(case + "_")}
    raw = out.pop("blocks")
    blocks = np.zeros(len(raw), dtype=N.block_dtype())
    for i, name in enumerate(("row_ofs", "col_ofs", "nrows", "ncols", "rank", "kind", "off_a", "off_b", "off_s")):
        blocks[name] = raw[:, i]
    out["blocks"] = blocks
    out["nrows"], out["ncols"] = (int(v) for v in out.pop("shape"))
    out["alpha"] = float(out["alpha"][0])
    return out


def exact_block(g: dict, b: int) -> np.ndarray:
    """the integer block b of the fixture as an FP64 array (S folded in)"""
    blk = g["blocks"][b]
    ar = g["arena"]
    m, n, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
    if int(blk["kind"]) == N.BLOCK_DENSE:
        return ar[blk["off_a"]:blk["off_a"] + m * n].reshape(n, m).T
    U = ar[blk["off_a"]:blk["off_a"] + m * k].reshape(k, m).T
    V = ar[blk["off_b"]:blk["off_b"] + n * k].reshape(k, n).T
    if int(blk["kind"]) == N.BLOCK_LOWRANK_SV:
        U = U * ar[blk["off_s"]:blk["off_s"] + k]
    return U @ V.T


def abs_bound(g: dict, x: np.ndarray, trans: bool) -> float:
    """alpha * max_i sum over all products of |entries| in row i of op(A) |x|:
    a bound on every partial sum any evaluation order can form"""
    acc = np.zeros(g["ncols"] if trans else g["nrows"])
    for b, blk in enumerate(g["blocks"]):
        r0, c0, m, n = int(blk["row_ofs"]), int(blk["col_ofs"]), int(blk["nrows"]), int(blk["ncols"])
        ar, k = g["arena"], int(blk["rank"])
        if int(blk["kind"]) == N.BLOCK_DENSE:
            D = np.abs(ar[blk["off_a"]:blk["off_a"] + m * n].reshape(n, m).T)
            if trans:
                acc[c0:c0 + n] += D.T @ np.abs(x[r0:r0 + m])
            else:
                acc[r0:r0 + m] += D @ np.abs(x[c0:c0 + n])
            continue
        U = np.abs(ar[blk["off_a"]:blk["off_a"] + m * k].reshape(k, m).T)
        V = np.abs(ar[blk["off_b"]:blk["off_b"] + n * k].reshape(k, n).T)
        S = ar[blk["off_s"]:blk["off_s"] + k] if int(blk["kind"]) == N.BLOCK_LOWRANK_SV else np.ones(k)
        if trans:
            acc[c0:c0 + n] += V @ ((U * S).T @ np.abs(x[r0:r0 + m]))
        else:
            acc[r0:r0 + m] += U @ ((V * S).T @ np.abs(x[c0:c0 + n]))
    return abs(g["alpha"]) * float(acc.max())
