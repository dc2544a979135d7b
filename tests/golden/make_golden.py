"""Writes exact_trees.npz: hand-built block trees with small-integer factors
and vectors, and their products computed in exact integer arithmetic.

The reference ships no golden vectors for its mul_vec path, so these
fixtures pin it by mathematics instead: every factor, singular value, x, y0
and alpha is a small integer, so every partial sum of y0 + alpha op(A) x is
an integer far below 2^53 and any correct FP64 evaluation, in any summation
order, returns the exact result bit for bit.  The expected outputs come from
Python integers (no floating point, no numpy arithmetic, no oracle).

Cases (all blocks in row-major recursion order, as the reference's
hlr::mul_vec visits them, arith/mulvec.hh:27-64):
  h_ragged   300 x 260 H-matrix from two ragged binary cluster trees with an
             interval admissibility rule; dense, UV and U diag(S) V^T leaves,
             ranks 0..6
  tlr        230 x 230 one-level tiling with tile 40 (last tile 30), dense
             diagonal tiles and low-rank off-diagonal tiles (tlr::mul_vec,
             arith/mulvec.hh:882-919)
For each case: normal and transposed products with alpha = 2 into a nonzero
y0, and a 3-column (multi right-hand side) normal product.

Run: python tests/golden/make_golden.py   (deterministic, seed fixed)
"""
from __future__ import annotations

import os
import random

import numpy as np

DENSE, LOWRANK, LOWRANK_SV = 0, 1, 2


def cluster(lo, hi, leaf, rng):
    """ragged binary tree over [lo, hi): list of (lo, hi, children)"""
    if hi - lo <= leaf:
        return (lo, hi, [])
    mid = (lo + hi) // 2 + rng.randint(-(hi - lo) // 8, (hi - lo) // 8)
    return (lo, hi, [cluster(lo, mid, leaf, rng), cluster(mid, hi, leaf, rng)])


def admissible(t, s):
    # intervals separated by at least the smaller length
    gap = max(s[0] - t[1], t[0] - s[1])
    return gap >= min(t[1] - t[0], s[1] - s[0])


def block_tree(t, s, out):
    """leaves of the block tree in row-major recursion order"""
    if admissible(t, s):
        out.append(("lr", t, s))
    elif not t[2] or not s[2]:
        out.append(("dense", t, s))
    else:
        for tc in t[2]:
            for sc in s[2]:
                block_tree(tc, sc, out)


class Builder:
    def __init__(self, rng):
        self.rng = rng
        self.arena: list[int] = []
        self.blocks: list[list[int]] = []
        self.leaves = []          # (kind, r0, c0, m, n, D | (U, V, S))

    def _put(self, vals):
        off = len(self.arena)
        self.arena.extend(vals)
        return off

    def dense(self, r0, c0, m, n):
        D = [[self.rng.randint(-5, 5) for _ in range(n)] for _ in range(m)]
        off = self._put([D[i][j] for j in range(n) for i in range(m)])      # column-major
        self.blocks.append([r0, c0, m, n, 0, DENSE, off, -1, -1])
        self.leaves.append((DENSE, r0, c0, m, n, D))

    def lowrank(self, r0, c0, m, n, k, sv):
        U = [[self.rng.randint(-3, 3) for _ in range(k)] for _ in range(m)]
        V = [[self.rng.randint(-3, 3) for _ in range(k)] for _ in range(n)]
        S = [self.rng.randint(1, 4) for _ in range(k)] if sv else [1] * k
        offu = self._put([U[i][l] for l in range(k) for i in range(m)])
        offv = self._put([V[j][l] for l in range(k) for j in range(n)])
        offs = self._put(S) if sv else -1
        self.blocks.append([r0, c0, m, n, k, LOWRANK_SV if sv else LOWRANK, offu, offv, offs])
        self.leaves.append((LOWRANK_SV if sv else LOWRANK, r0, c0, m, n, (U, V, S)))

    def entry(self, leaf, i, j):
        kind, r0, c0, m, n, F = leaf
        if kind == DENSE:
            return F[i][j]
        U, V, S = F
        return sum(U[i][l] * S[l] * V[j][l] for l in range(len(S)))

    def matrix(self, nrows, ncols):
        A = [[0] * ncols for _ in range(nrows)]
        for leaf in self.leaves:
            _, r0, c0, m, n, _ = leaf
            for i in range(m):
                for j in range(n):
                    A[r0 + i][c0 + j] += self.entry(leaf, i, j)
        return A


def products(A, rng, nrows, ncols):
    alpha = 2
    x = [rng.randint(-4, 4) for _ in range(ncols)]
    y0 = [rng.randint(-10, 10) for _ in range(nrows)]
    y = [y0[i] + alpha * sum(A[i][j] * x[j] for j in range(ncols)) for i in range(nrows)]
    xt = [rng.randint(-4, 4) for _ in range(nrows)]
    yt0 = [rng.randint(-10, 10) for _ in range(ncols)]
    yt = [yt0[j] + alpha * sum(A[i][j] * xt[i] for i in range(nrows)) for j in range(ncols)]
    X = [[rng.randint(-4, 4) for _ in range(ncols)] for _ in range(3)]
    Y0 = [[rng.randint(-10, 10) for _ in range(nrows)] for _ in range(3)]
    Y = [[Y0[r][i] + alpha * sum(A[i][j] * X[r][j] for j in range(ncols)) for i in range(nrows)]
         for r in range(3)]
    return dict(alpha=alpha, x=x, y0=y0, y=y, xt=xt, yt0=yt0, yt=yt, X=X, Y0=Y0, Y=Y)


def case_h(rng):
    nrows, ncols = 300, 260
    rt, ct = cluster(0, nrows, 48, rng), cluster(0, ncols, 48, rng)
    leaves = []
    block_tree(rt, ct, leaves)
    b = Builder(rng)
    for kind, t, s in leaves:
        m, n = t[1] - t[0], s[1] - s[0]
        if kind == "dense":
            b.dense(t[0], s[0], m, n)
        else:
            b.lowrank(t[0], s[0], m, n, rng.randint(0, 6), sv=rng.random() < 0.4)
    return b, nrows, ncols


def case_tlr(rng):
    n, tile = 230, 40
    edges = list(range(0, n, tile)) + [n]
    b = Builder(rng)
    for bi in range(len(edges) - 1):
        for bj in range(len(edges) - 1):
            r0, c0 = edges[bi], edges[bj]
            m, w = edges[bi + 1] - r0, edges[bj + 1] - c0
            if bi == bj:
                b.dense(r0, c0, m, w)
            else:
                b.lowrank(r0, c0, m, w, rng.randint(0, 5), sv=False)
    return b, n, n


def main():
    rng = random.Random(20240601)
    out = {}
    for name, fn in (("h_ragged", case_h), ("tlr", case_tlr)):
        b, nrows, ncols = fn(rng)
        A = b.matrix(nrows, ncols)
        p = products(A, rng, nrows, ncols)
        out[f"{name}_shape"] = np.array([nrows, ncols], dtype=np.int64)
        out[f"{name}_blocks"] = np.array(b.blocks, dtype=np.int64)
        out[f"{name}_arena"] = np.array(b.arena, dtype=np.float64)
        out[f"{name}_alpha"] = np.array([p["alpha"]], dtype=np.float64)
        for key in ("x", "y0", "y", "xt", "yt0", "yt", "X", "Y0", "Y"):
            out[f"{name}_{key}"] = np.array(p[key], dtype=np.float64)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "exact_trees.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: " + ", ".join(f"{k} {v.shape}" for k, v in out.items() if k.endswith("_blocks")))


if __name__ == "__main__":
    main()
