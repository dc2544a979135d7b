"""Uniform-H and H^2 matrices: cluster bases shared by all blocks of a block
row (column), coupling matrices per admissible block, and the device plan
over the C-ABI (hmvm_uplan_*).

`build_uniform(hm, eps, nested)` converts a generated H-matrix into either
form.  It restates the reference's conversion in its own construction
(seq/detail/uniform_matrix.hh:369-530 extends row and column bases block by
block; here each basis is computed once from all blocks of its block row):

  shared bases   U_t = leading left singular vectors of [U_b L_b]_b over the
                 low-rank leaves b of block row t, with L_b L_b^T = V_b^T V_b
                 (so that U_b L_b spans the block's column space with its
                 weight), cut where the singular values fall below
                 eps * the largest; the same for the column bases;
  nested bases   bottom-up over the cluster tree: a leaf cluster takes the
                 same construction over the blocks of its own block row and
                 of every ancestor's (restricted to its indices); a parent
                 cluster works in the coordinates of its sons' bases, and
                 its leading eigenvectors give the transfer matrices E_s
                 (arith/detail/h2.hh:238-286 reads them as transfer_mat(s));
  couplings      S_b = (U_t^T U_b) (V_s^T V_b)^T for the explicit or
                 expanded bases U_t, V_s of the block's clusters.

The singular values of each basis go with it (off_s) for the per-column
storage chooser.  Dense leaves are kept as they are.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .plan import PRECISIONS, _stream_handle


@dataclass
class UniformHMatrix:
    rowcb: np.ndarray      # hmvm_cluster records
    colcb: np.ndarray
    blocks: np.ndarray     # hmvm_ublock records
    arena: np.ndarray      # float64
    nrows: int
    ncols: int
    nested: bool

    @property
    def n(self) -> int:
        return self.nrows


def _factors(hm, b):
    m, n, k = int(b["nrows"]), int(b["ncols"]), int(b["rank"])
    a = hm.arena
    U = a[b["off_a"]:b["off_a"] + m * k].reshape(k, m).T
    V = a[b["off_b"]:b["off_b"] + n * k].reshape(k, n).T
    if int(b["kind"]) == N.BLOCK_LOWRANK_SV:
        U = U * a[b["off_s"]:b["off_s"] + k]
    return U, V


def _weight(V):
    """L with L L^T = V^T V (eigen square root; V may be rank deficient)"""
    w, Q = np.linalg.eigh(V.T @ V)
    return Q * np.sqrt(np.clip(w, 0.0, None))


def _cut(s, eps):
    if s.size == 0 or s[0] <= 0:
        return 0
    return int((s > eps * s[0]).sum())


def _tree(intervals):
    """parent of each interval (ofs, size) in a laminar family: the smallest
    interval strictly containing it, or -1"""
    order = sorted(range(len(intervals)), key=lambda i: (intervals[i][0], -intervals[i][1]))
    parent = [-1] * len(intervals)
    stack = []
    for i in order:
        o, s = intervals[i]
        while stack and not (intervals[stack[-1]][0] <= o and o + s <= intervals[stack[-1]][0] + intervals[stack[-1]][1]):
            stack.pop()
        parent[i] = stack[-1] if stack else -1
        stack.append(i)
    return parent


class _Side:
    """clusters of one side (rows or columns) and their low-rank blocks"""

    def __init__(self, hm, rows: bool):
        key_o, key_s = ("row_ofs", "nrows") if rows else ("col_ofs", "ncols")
        ivs = {}
        for blk in hm.blocks:
            ivs.setdefault((int(blk[key_o]), int(blk[key_s])), None)
        self.intervals = sorted(ivs)
        self.index = {iv: i for i, iv in enumerate(self.intervals)}
        self.parent = _tree(self.intervals)
        self.blocks = [[] for _ in self.intervals]      # (leaf index, factor on this side, weight)
        for bi, blk in enumerate(hm.blocks):
            if int(blk["kind"]) == N.BLOCK_DENSE or int(blk["rank"]) == 0:
                continue
            U, V = _factors(hm, blk)
            own, other = (U, V) if rows else (V, U)
            self.blocks[self.index[(int(blk[key_o]), int(blk[key_s]))]].append((bi, own, _weight(other)))
        self.sons = [[] for _ in self.intervals]
        for i, p in enumerate(self.parent):
            if p >= 0:
                self.sons[p].append(i)

    def shared(self, eps):
        """explicit basis (|t| x k) and singular values per cluster"""
        out = []
        for t, (o, s) in enumerate(self.intervals):
            if not self.blocks[t]:
                out.append((np.zeros((s, 0)), np.zeros(0)))
                continue
            W = np.hstack([f @ w for _, f, w in self.blocks[t]])
            Q, sv, _ = np.linalg.svd(W, full_matrices=False)
            k = _cut(sv, eps)
            out.append((Q[:, :k].copy(), sv[:k].copy()))
        return out

    def nested(self, eps):
        """per cluster: explicit basis (leaves) or transfer E (rank x parent
        rank, for sons), singular values, and the expanded basis"""
        nc = len(self.intervals)
        depth = [0] * nc
        for i in range(nc):
            p = self.parent[i]
            while p >= 0:
                depth[i] += 1
                p = self.parent[p]
        # blocks acting on each cluster: its own and every ancestor's
        acting = [[] for _ in range(nc)]
        for t in range(nc):
            a = t
            while a >= 0:
                acting[t].extend((a, f, w) for _, f, w in self.blocks[a])
                a = self.parent[a]
        basis, sval, coords = [None] * nc, [None] * nc, [None] * nc
        for t in sorted(range(nc), key=lambda i: -depth[i]):
            o, s = self.intervals[t]
            if not self.sons[t]:
                # leaf: W = [F_b(rows of t) L_b]
                if not acting[t]:
                    basis[t], sval[t] = np.zeros((s, 0)), np.zeros(0)
                    continue
                W = np.hstack([f[o - self.intervals[a][0]:o - self.intervals[a][0] + s] @ w for a, f, w in acting[t]])
                Q, sv, _ = np.linalg.svd(W, full_matrices=False)
                k = _cut(sv, eps)
                basis[t], sval[t] = Q[:, :k].copy(), sv[:k].copy()
                continue
            # parent: coordinates of the acting blocks in the sons' bases
            parts = []
            for a, f, w in acting[t]:
                ao = self.intervals[a][0]
                rows = []
                for c in self.sons[t]:
                    co, cs = self.intervals[c]
                    rows.append(basis_expanded(self, basis, coords, c).T @ f[co - ao:co - ao + cs])
                parts.append(np.vstack(rows) @ w)
            if not parts or sum(basis_rank(basis, coords, c) for c in self.sons[t]) == 0:
                basis[t], sval[t] = None, np.zeros(0)
                coords[t] = np.zeros((0, 0))
                continue
            W = np.hstack(parts)
            Q, sv, _ = np.linalg.svd(W, full_matrices=False)
            k = _cut(sv, eps)
            coords[t], sval[t] = Q[:, :k].copy(), sv[:k].copy()
        return basis, sval, coords


def basis_rank(basis, coords, c):
    return basis[c].shape[1] if basis[c] is not None else coords[c].shape[1]


def basis_expanded(side, basis, coords, c):
    """expanded basis of cluster c (|c| x k): explicit, or the sons'
    expanded bases times their rows of the transfer coordinates"""
    if basis[c] is not None:
        return basis[c]
    o, s = side.intervals[c]
    k = coords[c].shape[1]
    out = np.zeros((s, k))
    r0 = 0
    for son in side.sons[c]:
        so, ss = side.intervals[son]
        ks = basis_rank(basis, coords, son)
        out[so - o:so - o + ss] = basis_expanded(side, basis, coords, son) @ coords[c][r0:r0 + ks]
        r0 += ks
    return out


def build_uniform(hm, eps: float = 1e-6, nested: bool = False) -> UniformHMatrix:
    """uniform-H (shared bases) or H^2 (nested bases) form of H-matrix hm"""
    sides = [_Side(hm, True), _Side(hm, False)]
    parts, pos = [], 0

    def put(a) -> int:
        nonlocal pos
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        parts.append(a)
        off = pos
        pos += a.size
        return off

    records, expanded = [], []
    for side in sides:
        nc = len(side.intervals)
        rec = np.zeros(nc, dtype=N.cluster_dtype())
        rec["parent"] = -1
        rec["off_basis"] = rec["off_transfer"] = rec["off_s"] = -1
        exp = [None] * nc
        if not nested:
            for t, (B, sv) in enumerate(side.shared(eps)):
                o, s = side.intervals[t]
                rec[t]["ofs"], rec[t]["size"], rec[t]["rank"] = o, s, B.shape[1]
                if B.shape[1]:
                    rec[t]["off_basis"] = put(B.T)
                    rec[t]["off_s"] = put(sv)
                exp[t] = B
        else:
            basis, sval, coords = side.nested(eps)
            for t in range(nc):
                o, s = side.intervals[t]
                k = basis_rank(basis, coords, t)
                rec[t]["ofs"], rec[t]["size"], rec[t]["rank"] = o, s, k
                rec[t]["parent"] = side.parent[t]
                if basis[t] is not None and k:
                    rec[t]["off_basis"] = put(basis[t].T)
                if k:
                    rec[t]["off_s"] = put(sval[t])
                exp[t] = basis_expanded(side, basis, coords, t)
            # transfer of each son: its rows of the parent's coordinates
            for p in range(nc):
                if basis[p] is not None or rec[p]["rank"] == 0:
                    continue
                r0 = 0
                for son in side.sons[p]:
                    ks = int(rec[son]["rank"])
                    if ks:
                        rec[son]["off_transfer"] = put(coords[p][r0:r0 + ks].T)   # ks x kp, column-major
                    r0 += ks
        records.append(rec)
        expanded.append(exp)
    rowcb, colcb = records
    ublocks = np.zeros(len(hm.blocks), dtype=N.ublock_dtype())
    for bi, blk in enumerate(hm.blocks):
        ub = ublocks[bi]
        ub["row_ofs"], ub["col_ofs"] = blk["row_ofs"], blk["col_ofs"]
        ub["nrows"], ub["ncols"] = blk["nrows"], blk["ncols"]
        m, n = int(blk["nrows"]), int(blk["ncols"])
        if int(blk["kind"]) == N.BLOCK_DENSE:
            ub["kind"] = N.BLOCK_DENSE
            ub["rowcb"] = ub["colcb"] = -1
            ub["off"] = put(hm.arena[blk["off_a"]:blk["off_a"] + m * n])
            continue
        r = sides[0].index[(int(blk["row_ofs"]), m)]
        c = sides[1].index[(int(blk["col_ofs"]), n)]
        ub["kind"], ub["rowcb"], ub["colcb"] = N.BLOCK_COUPLING, r, c
        Ur, Vc = expanded[0][r], expanded[1][c]
        if int(blk["rank"]) == 0 or Ur.shape[1] == 0 or Vc.shape[1] == 0:
            S = np.zeros((Ur.shape[1], Vc.shape[1]))
        else:
            U, V = _factors(hm, blk)
            S = (Ur.T @ U) @ (Vc.T @ V).T
        ub["off"] = put(S.T)
    arena = np.concatenate(parts) if parts else np.zeros(0)
    return UniformHMatrix(rowcb, colcb, ublocks, arena, hm.nrows, hm.ncols, nested)


class UPlan:
    """hmvm_uplan wrapper: device plan of a uniform-H / H^2 matrix."""

    def __init__(self, um: UniformHMatrix, precision: str = "fp64", eps: float = 0.0, device: int = 0):
        self._p = None
        opts = N.HmvmPlanOpts()
        opts.precision = PRECISIONS[precision]
        opts.dense_prec = N.PREC_VALR
        opts.eps = float(eps)
        opts.check_finite = 1
        opts.force_format = 0

        def ptr(a):
            return C.c_void_p(a.ctypes.data) if a.size else None
        p = C.c_void_p()
        N.check(N.hmvm().hmvm_uplan_create(ptr(um.rowcb), len(um.rowcb), ptr(um.colcb), len(um.colcb),
                                           ptr(um.blocks), len(um.blocks), um.nrows, um.ncols, ptr(um.arena),
                                           um.arena.size, C.byref(opts), device, C.byref(p)), "hmvm_uplan_create")
        self._p = p
        self.device = device
        self.nrows, self.ncols = um.nrows, um.ncols

    def close(self) -> None:
        if self._p:
            N.hmvm().hmvm_uplan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = N.HmvmUplanInfo()
        N.check(N.hmvm().hmvm_uplan_get_info(self._p, C.byref(i)), "hmvm_uplan_get_info")
        return {name: getattr(i, name) for name, _ in N.HmvmUplanInfo._fields_}

    def apply(self, alpha: float, op: int, x, y, stream=None) -> None:
        """y += alpha op(M) x for float64 device vectors of the plan's sizes"""
        import torch
        trans = op in (N.OP_TRANSPOSED, N.OP_ADJOINT)
        nin, nout = (self.nrows, self.ncols) if trans else (self.ncols, self.nrows)
        for t, name, n in ((x, "x", nin), (y, "y", nout)):
            if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or t.dim() != 1 or t.numel() != n:
                raise ValueError(f"uplan apply: {name} must be a 1-D float64 tensor of {n} entries")
            if n > 1 and t.stride(0) != 1:
                raise ValueError(f"uplan apply: {name} must have unit stride")
            if t.device.type != "cuda" or t.device.index != self.device:
                raise ValueError(f"uplan apply: {name} is on {t.device}, the plan on HIP device {self.device}")
        N.check(N.hmvm().hmvm_uplan_apply(self._p, float(alpha), int(op), C.c_void_p(x.data_ptr()),
                                          C.c_void_p(y.data_ptr()), _stream_handle(stream)), "hmvm_uplan_apply")
