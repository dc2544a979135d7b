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
from .plan import Plan, _stream_handle


class Batch:
    def __init__(self, plans: list[Plan], nrhs: list[int], vec_dtype: int = N.F64):
        self._b = None
        self.plans = list(plans)
        self.nrhs = [int(r) for r in nrhs]
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
        al = (C.c_double * len(xs))(*[float(a) for a in (alphas or [1.0] * len(xs))])
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
