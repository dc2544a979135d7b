"""Batches of H-matrix-vector products (libhmvm.so).

`Batch(plans, nrhs)` collects forward products y_i += alpha_i A_i x_i of many
plans with per-member numbers of right-hand sides; `apply` runs the work of
all members as at most twelve stream-ordered kernels (one unit of work per
workgroup),
`capture`/`replay` record that sequence in a HIP graph, `tune` runs the
in-project autotuner over unit order and partial-sum granularity.
"""
from __future__ import annotations

import ctypes as C

from . import _native as N
from .plan import Plan, _stream_handle, _vec_dtype, check_vectors


class Batch:
    def __init__(self, plans: list[Plan], nrhs: list[int], vec_dtype: int = N.F64):
        self._b = None
        self.plans = list(plans)
        self.nrhs = [int(r) for r in nrhs]
        if len(self.nrhs) != len(self.plans):
            raise ValueError(f"batch: {len(self.plans)} plans but {len(self.nrhs)} right-hand-side counts")
        self.vec_dtype = int(vec_dtype)
        arr = (C.c_void_p * len(plans))(*[p.handle for p in plans])
        rr = (C.c_int * len(plans))(*self.nrhs)
        b = C.c_void_p()
        N.check(N.hmvm().hmvm_batch_create(arr, len(plans), rr, vec_dtype, C.byref(b)), "hmvm_batch_create")
        self._b = b

    def close(self) -> None:
        if self._b:
            N.hmvm().hmvm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _vecs(tensors):
        ptrs = (C.c_void_p * len(tensors))(*[C.c_void_p(t.data_ptr()) for t in tensors])
        lds = (C.c_int64 * len(tensors))(*[(t.stride(0) if t.dim() == 2 else t.numel()) for t in tensors])
        return ptrs, lds

    def _args(self, alphas, xs, ys):
        """Checked ctypes arrays of the members' alphas and vectors: one x, y
        (and alpha) per member, each pair valid for its plan with exactly the
        member's number of right-hand sides (the C side reads nplans entries
        of every array and cannot check the tensors)."""
        xs, ys = list(xs), list(ys)
        n = len(self.plans)
        if len(xs) != n or len(ys) != n:
            raise ValueError(f"batch: {n} members, got {len(xs)} x and {len(ys)} y tensors")
        alphas = [1.0] * n if alphas is None else list(alphas)
        if len(alphas) != n:
            raise ValueError(f"batch: {n} members, got {len(alphas)} alphas")
        for i, (p, r, x, y) in enumerate(zip(self.plans, self.nrhs, xs, ys)):
            check_vectors(p, N.OP_NORMAL, x, y, f"batch member {i}", nrhs=r, device=False)
            if r > 1 and x.dim() != 2:
                raise ValueError(f"batch member {i}: {r} right-hand sides need (nrhs, n) tensors")
            if _vec_dtype(x) != self.vec_dtype:
                raise ValueError(f"batch member {i}: vectors of {x.dtype}, the batch was created for "
                                 f"{'float64' if self.vec_dtype == N.F64 else 'float32'}")
        for i, (p, x, y) in enumerate(zip(self.plans, xs, ys)):
            check_vectors(p, N.OP_NORMAL, x, y, f"batch member {i}")
        al = (C.c_double * n)(*[float(a) for a in alphas])
        xp, lx = self._vecs(xs)
        yp, ly = self._vecs(ys)
        return al, xp, lx, yp, ly

    def apply(self, xs, ys, alphas=None, stream=None) -> None:
        """xs[i], ys[i]: torch tensors (n,) or (nrhs, n) on the plans' device."""
        al, xp, lx, yp, ly = self._args(alphas, xs, ys)
        N.check(N.hmvm().hmvm_batch_apply(self._b, al, xp, lx, yp, ly, _stream_handle(stream)), "hmvm_batch_apply")

    def capture(self, xs, ys, alphas=None, stream=None) -> None:
        al, xp, lx, yp, ly = self._args(alphas, xs, ys)
        N.check(N.hmvm().hmvm_batch_capture(self._b, al, xp, lx, yp, ly, _stream_handle(stream)),
                "hmvm_batch_capture")

    def replay(self, stream=None) -> None:
        N.check(N.hmvm().hmvm_batch_replay(self._b, _stream_handle(stream)), "hmvm_batch_replay")

    def tune(self, xs, ys, stream=None) -> int:
        _, xp, lx, yp, ly = self._args(None, xs, ys)
        ch = C.c_int()
        N.check(N.hmvm().hmvm_batch_tune(self._b, xp, lx, yp, ly, _stream_handle(stream), C.byref(ch)),
                "hmvm_batch_tune")
        return int(ch.value)

    def check(self, stream=None) -> None:
        N.check(N.hmvm().hmvm_batch_check(self._b, _stream_handle(stream)), "hmvm_batch_check")
