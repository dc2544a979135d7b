"""Shared helpers of the test-suite (not product code)."""
from __future__ import annotations

import numpy as np

from hlr_amd import _native as N


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def dequantized_arena(plan, hm):
    """FP64 arena holding exactly the values the plan stores (dequantised),
    at the generator's offsets, so the oracle can run on them."""
    ar = [plan.arena_numpy(f) for f in range(N.NFMT)]
    out = np.array(hm.arena, copy=True)
    for b, blk in enumerate(hm.blocks):
        m, n, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
        U, V = plan.unpack_block(b, ar)
        if int(blk["kind"]) == N.BLOCK_DENSE:
            out[blk["off_a"]:blk["off_a"] + m * n] = U.T.reshape(-1)
        elif k:
            out[blk["off_a"]:blk["off_a"] + m * k] = U.T.reshape(-1)
            out[blk["off_b"]:blk["off_b"] + n * k] = V.T.reshape(-1)
    return out
