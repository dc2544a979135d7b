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


def inverse_blocks(hm, A: np.ndarray, tau: float, omega: float = 1.0):
    """An approximate inverse omega * A^-1 in hm's block structure, as a
    namespace with blocks / arena / n (test input for the consumers that
    take a preconditioner): dense leaves hold the exact blocks of the
    inverse, admissible leaves its truncated SVD U S V^T (LOWRANK_SV) with
    the singular values above tau |A^-1|_2."""
    from types import SimpleNamespace
    Ai = omega * np.linalg.inv(A)
    thr = tau * np.linalg.norm(Ai, 2)
    blocks = np.array(hm.blocks, copy=True)
    parts, off = [], 0
    for blk in blocks:
        r0, c0, m, n = int(blk["row_ofs"]), int(blk["col_ofs"]), int(blk["nrows"]), int(blk["ncols"])
        S = Ai[r0:r0 + m, c0:c0 + n]
        if int(blk["kind"]) == N.BLOCK_DENSE:
            blk["off_a"] = off
            parts.append(S.T.reshape(-1))
            off += m * n
            continue
        U, s, Vt = np.linalg.svd(S, full_matrices=False)
        k = int((s > thr).sum())
        blk["kind"], blk["rank"] = N.BLOCK_LOWRANK_SV, k
        blk["off_a"] = off
        parts.append(U[:, :k].T.reshape(-1))
        off += m * k
        blk["off_b"] = off
        parts.append(Vt[:k].reshape(-1))
        off += n * k
        blk["off_s"] = off
        parts.append(s[:k])
        off += k
    return SimpleNamespace(blocks=blocks, arena=np.concatenate(parts), n=hm.n, nrows=hm.n, ncols=hm.n)
