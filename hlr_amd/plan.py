"""Device plans over the C-ABI (libhmvm.so).

A `Plan` owns the device layout of one H-/TLR-matrix: factor arenas in the
chosen storage precision plus the panel/index tables.  Vectors are torch
tensors on the plan's device; torch is used for device memory and streams
only, all arithmetic runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

PRECISIONS = {"fp64": N.PREC_FP64, "fp32": N.PREC_FP32, "valr": N.PREC_VALR, "bfp": N.PREC_BFP, "afl": N.PREC_AFL}
FORMAT_NAMES = ("fp64", "fp32", "bf16", "fp8_e4m3", "bfp8", "bfp16", "bfp25", "afl12", "afl16", "afl24")
# bytes per stored value (16 bytes hold 128 // bits values)
FORMAT_BYTES = tuple(16.0 / (128 // b) for b in N.FMT_BITS)


def _stream_handle(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(int(stream.cuda_stream))


def _vec_dtype(t):
    import torch
    if t.dtype == torch.float64:
        return N.F64
    if t.dtype == torch.float32:
        return N.F32
    raise TypeError(f"unsupported vector dtype {t.dtype}")


def _on_device(t, device: int, what: str) -> None:
    if t.device.type != "cuda" or t.device.index != device:
        raise ValueError(f"{what}: tensor on {t.device}, the plan lives on HIP device {device}")


def check_vectors(plan, op: int, x, y, what: str = "apply", nrhs: int | None = None, device: bool = True):
    """Validate the vectors of y += alpha op(A) x before any C call (the
    C-ABI takes raw pointers and cannot check sizes itself).

    x and y: torch tensors of one floating dtype (float64 or float32), either
    1-D of ncols(op) / nrows(op) entries with unit stride, or 2-D (nrhs, n)
    with unit stride along n and row stride >= n (column-major n x nrhs), on
    the plan's device (checked last, and only when `device` is true).  nrhs,
    when given, is the count the caller expects.  Returns (nrhs, ldx, ldy);
    raises TypeError / ValueError."""
    import torch
    for t, name in ((x, "x"), (y, "y")):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{what}: {name} is not a torch tensor")
    _vec_dtype(x)
    if y.dtype != x.dtype:
        raise ValueError(f"{what}: x is {x.dtype} but y is {y.dtype}")
    trans = op in (N.OP_TRANSPOSED, N.OP_ADJOINT)
    nin = plan.nrows if trans else plan.ncols
    nout = plan.ncols if trans else plan.nrows
    if x.dim() != y.dim() or x.dim() not in (1, 2):
        raise ValueError(f"{what}: x and y must both be 1-D or both (nrhs, n)")
    if x.dim() == 1:
        if x.numel() != nin or y.numel() != nout:
            raise ValueError(f"{what}: expected x of {nin} and y of {nout} entries, got {x.numel()} and {y.numel()}")
        if (nin > 1 and x.stride(0) != 1) or (nout > 1 and y.stride(0) != 1):
            raise ValueError(f"{what}: vectors must have unit stride")
        r, ldx, ldy = 1, 0, 0
    else:
        r = x.shape[0]
        if r < 1 or y.shape[0] != r:
            raise ValueError(f"{what}: x and y must hold the same number (>= 1) of right-hand sides")
        if x.shape[1] != nin or y.shape[1] != nout:
            raise ValueError(f"{what}: expected (nrhs, {nin}) and (nrhs, {nout}), got {tuple(x.shape)} and "
                             f"{tuple(y.shape)}")
        if (nin > 1 and x.stride(1) != 1) or (nout > 1 and y.stride(1) != 1):
            raise ValueError(f"{what}: x, y must be (nrhs, n) tensors with unit stride along n")
        ldx, ldy = x.stride(0), y.stride(0)
        if r > 1 and (ldx < nin or ldy < nout):
            raise ValueError(f"{what}: row stride below the vector length")
    if nrhs is not None and r != nrhs:
        raise ValueError(f"{what}: expected {nrhs} right-hand sides, got {r}")
    if device:
        _on_device(x, plan.device, what)
        _on_device(y, plan.device, what)
    return r, ldx, ldy


class Plan:
    """hmvm_plan wrapper.

    blocks: numpy array of hmvm_block records, arena: float64 numpy array.
    device: HIP device ordinal, or -1 for a host-only layout (no apply).
    """

    def __init__(self, blocks: np.ndarray, arena: np.ndarray, nrows: int, ncols: int,
                 precision: str = "fp64", eps: float = 0.0, dense_prec: str = "auto",
                 device: int = 0, check_finite: bool = True, force_format: str | None = None):
        self._p = None
        blocks = np.ascontiguousarray(blocks, dtype=N.block_dtype())
        arena = np.ascontiguousarray(arena, dtype=np.float64)
        opts = N.HmvmPlanOpts()
        opts.precision = PRECISIONS[precision]
        # dense leaves under "valr" / "bfp" / "afl": "fp64", "fp32", or "auto"
        # (valr: FP32 while eps needs at most 23 mantissa bits, else FP64;
        # bfp / afl: the plan's packed-format column rule at rate(eps))
        opts.dense_prec = {"fp64": N.PREC_FP64, "fp32": N.PREC_FP32, "auto": N.PREC_VALR}[dense_prec]
        opts.eps = float(eps)
        opts.check_finite = 1 if check_finite else 0
        opts.force_format = 0 if force_format is None else 1 + FORMAT_NAMES.index(force_format)
        p = C.c_void_p()
        rc = N.hmvm().hmvm_plan_create(blocks.ctypes.data if len(blocks) else None, len(blocks), nrows, ncols,
                                       arena.ctypes.data if arena.size else None, arena.size, C.byref(opts),
                                       device, C.byref(p))
        N.check(rc, "hmvm_plan_create")
        self._p = p
        self.device = device
        self.precision = precision
        self.nrows, self.ncols = nrows, ncols
        self.blocks = blocks

    @classmethod
    def from_hmatrix(cls, hm, **kw) -> "Plan":
        return cls(hm.blocks, hm.arena, hm.nrows, hm.ncols, **kw)

    @classmethod
    def from_file(cls, path, **kw) -> "Plan":
        """Plan of the matrix in an H-matrix file, in any storage precision
        (the file is read into host memory, then the plan is built from it;
        same options as the constructor)."""
        from .hmatrix import load_hmatrix
        hm = load_hmatrix(path)
        return cls(hm.blocks, hm.arena, hm.nrows, hm.ncols, **kw)

    def close(self) -> None:
        if self._p:
            N.hmvm().hmvm_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._p

    def info(self) -> dict:
        i = N.HmvmPlanInfo()
        N.check(N.hmvm().hmvm_plan_get_info(self._p, C.byref(i)), "hmvm_plan_get_info")
        d = {name: getattr(i, name) for name, _ in N.HmvmPlanInfo._fields_ if name != "reserved"}
        d["bytes_by_format"] = list(i.bytes_by_format)
        d["cols_by_format"] = list(i.cols_by_format)
        return d

    # -- products -----------------------------------------------------------
    def apply(self, alpha: float, op: int, x, y, stream=None) -> None:
        """y += alpha op(A) x for torch tensors (1-D, or 2-D column-major
        n x nrhs, i.e. a transposed contiguous nrhs x n tensor)."""
        # column-major n x nrhs == row-major nrhs x n
        nrhs, ldx, ldy = check_vectors(self, op, x, y, "apply")
        rc = N.hmvm().hmvm_apply(self._p, float(alpha), int(op), C.c_void_p(x.data_ptr()), ldx,
                                 C.c_void_p(y.data_ptr()), ldy, nrhs, _vec_dtype(x), _stream_handle(stream))
        N.check(rc, "hmvm_apply")

    def apply_timed(self, alpha: float, op: int, x, y, stream=None):
        """apply with HIP events around the kernel phases; returns
        (reduction pass, partial sums, output pass) durations in ms."""
        if x.dim() != 1:
            raise ValueError("apply_timed: one right-hand side, 1-D vectors")
        check_vectors(self, op, x, y, "apply_timed", nrhs=1)
        ms = (C.c_float * 3)()
        rc = N.hmvm().hmvm_apply_timed(self._p, float(alpha), int(op), C.c_void_p(x.data_ptr()),
                                       C.c_void_p(y.data_ptr()), _vec_dtype(x), _stream_handle(stream), ms)
        N.check(rc, "hmvm_apply_timed")
        return tuple(float(v) for v in ms)

    def apply_ws(self, alpha, op, x, y, ws, stream=None) -> None:
        """apply with a caller workspace `ws` (a tensor on the plan's device
        of at least workspace_bytes(op, nrhs) bytes)."""
        nrhs, ldx, ldy = check_vectors(self, op, x, y, "apply_ws")
        _check_buffer(ws, self.device, self.workspace_bytes(op, nrhs), "apply_ws: workspace")
        rc = N.hmvm().hmvm_apply_ws(self._p, float(alpha), int(op), C.c_void_p(x.data_ptr()), ldx,
                                    C.c_void_p(y.data_ptr()), ldy, nrhs, _vec_dtype(x), C.c_void_p(ws.data_ptr()),
                                    ws.numel() * ws.element_size(), _stream_handle(stream))
        N.check(rc, "hmvm_apply_ws")

    def workspace_bytes(self, op: int = N.OP_NORMAL, nrhs: int = 1) -> int:
        b = C.c_int64()
        N.check(N.hmvm().hmvm_workspace_size(self._p, op, nrhs, C.byref(b)), "hmvm_workspace_size")
        return int(b.value)

    def tlr_apply(self, alpha: float, op: int, x, y, stream=None) -> None:
        if x.dim() != 1:
            raise ValueError("tlr_apply: one right-hand side, 1-D vectors")
        check_vectors(self, op, x, y, "tlr_apply", nrhs=1)
        rc = N.hmvm().hmvm_tlr_apply(self._p, float(alpha), int(op), C.c_void_p(x.data_ptr()),
                                     C.c_void_p(y.data_ptr()), _vec_dtype(x), _stream_handle(stream))
        N.check(rc, "hmvm_tlr_apply")

    def spectral_norm(self, atol: float = 0.0, max_it: int = 0, squared: bool = True, stream=None):
        """Spectral norm by power iteration on the device (the reference's
        hlr::norm::spectral, arith/detail/norm.hh:448-527): returns
        (lambda, iterations).  atol = 0: sqrt(eps); max_it = 0:
        max(5, min(nrows, ncols) / 10)."""
        lam = C.c_double()
        its = C.c_int64()
        N.check(N.hmvm().hmvm_spectral_norm(self._p, float(atol), int(max_it), 1 if squared else 0,
                                            _stream_handle(stream), C.byref(lam), C.byref(its)),
                "hmvm_spectral_norm")
        return float(lam.value), int(its.value)

    # -- arenas and gradients -------------------------------------------------
    def arena_elems(self, fmt: int) -> int:
        e = C.c_int64()
        N.check(N.hmvm().hmvm_plan_arena_elems(self._p, fmt, C.byref(e)), "hmvm_plan_arena_elems")
        return int(e.value)

    def arena_bytes(self, fmt: int) -> int:
        return int(self.info()["bytes_by_format"][fmt])

    def arena_numpy(self, fmt: int) -> np.ndarray:
        """Host copy of the stored arena of format fmt (host-only or device
        plan); block floating point arenas come as their raw bytes."""
        dt = (np.float64, np.float32, np.uint16, np.uint8)[fmt] if fmt < N.FMT_B8 else np.uint8
        out = np.zeros(self.arena_bytes(fmt) // np.dtype(dt).itemsize, dtype=dt)
        if out.size == 0:
            return out
        if self.device < 0:
            N.check(N.hmvm().hmvm_plan_get_arena(self._p, fmt, C.c_void_p(out.ctypes.data), None), "get_arena")
            return out
        import torch
        t = torch.empty(out.nbytes, dtype=torch.uint8, device=f"cuda:{self.device}")
        N.check(N.hmvm().hmvm_plan_get_arena(self._p, fmt, C.c_void_p(t.data_ptr()), _stream_handle(None)),
                "get_arena")
        out[...] = t.cpu().numpy().view(dt)
        return out

    def set_arena(self, fmt: int, src) -> None:
        """Replace the stored arena of format fmt by the device tensor src
        (at least arena_bytes(fmt) bytes, in the plan's layout)."""
        _check_buffer(src, self.device, self.arena_bytes(fmt), "set_arena")
        N.check(N.hmvm().hmvm_plan_set_arena(self._p, fmt, C.c_void_p(src.data_ptr()), _stream_handle(None)),
                "set_arena")

    def unpack_block(self, b: int, arenas, as_f64: bool = False):
        """Dequantised (U, V) or (D, None) of block b from host arenas."""
        blk = self.blocks[b]
        m, n, k = int(blk["nrows"]), int(blk["ncols"]), int(blk["rank"])
        dense = int(blk["kind"]) == N.BLOCK_DENSE
        a = np.zeros((m * n) if dense else (m * k))
        bb = None if dense else np.zeros(n * k)
        arenas = list(arenas) + [None] * (N.NFMT - len(arenas))
        ptrs = (C.c_void_p * N.NFMT)(*[C.c_void_p(arr.ctypes.data) if arr is not None and arr.size else None
                                       for arr in arenas])
        N.check(N.hmvm().hmvm_plan_unpack_block(self._p, b, ptrs, 1 if as_f64 else 0, C.c_void_p(a.ctypes.data),
                                                C.c_void_p(bb.ctypes.data) if bb is not None else None),
                "hmvm_plan_unpack_block")
        if dense:
            return a.reshape(n, m).T, None
        return a.reshape(k, m).T, bb.reshape(k, n).T

    def arena_source(self, fmt: int) -> np.ndarray:
        """int64 array, one entry per stored value of format fmt: the offset
        of that value in the FP64 arena the plan was built from (the leaves'
        column-major factors), or -1 for padding."""
        n = self.arena_elems(fmt)
        src = np.empty(n, dtype=np.int64)
        N.check(N.hmvm().hmvm_plan_arena_source(self._p, fmt, C.c_void_p(src.ctypes.data) if n else None),
                "hmvm_plan_arena_source")
        return src

    def grad(self, x, ybar, gx=None, factors: bool = True, stream=None):
        """Adjoint of y = A x: gx += A^T ybar, and per-format FP64 gradient
        arenas in the plan's layout.  For FP64, FP32 and BF16 storage an
        entry is dL/d(stored value); for the scaled formats (FP8 and block
        floating point) it is dL/d(dequantised value), i.e. with respect to
        the stored value times its panel-column scale."""
        import torch
        if x.dim() != 1 or ybar.dim() != 1:
            raise ValueError("grad: x and ybar must be 1-D")
        if gx is not None and (not isinstance(gx, torch.Tensor) or gx.dim() != 1 or gx.numel() != self.ncols
                               or gx.dtype != x.dtype or (self.ncols > 1 and gx.stride(0) != 1)):
            raise ValueError(f"grad: gx must be a unit-stride {x.dtype} vector of {self.ncols} entries")
        check_vectors(self, N.OP_NORMAL, x, ybar, "grad")
        if gx is not None:
            _on_device(gx, self.device, "grad (gx)")
        garena = [None] * N.NFMT
        ptrs = (C.c_void_p * N.NFMT)()
        if factors:
            for f in range(N.NFMT):
                e = self.arena_elems(f)
                if e:
                    garena[f] = torch.zeros(e, dtype=torch.float64, device=x.device)
                    ptrs[f] = C.c_void_p(garena[f].data_ptr())
        rc = N.hmvm().hmvm_grad(self._p, C.c_void_p(x.data_ptr()), C.c_void_p(ybar.data_ptr()),
                                C.c_void_p(gx.data_ptr()) if gx is not None else None,
                                ptrs if factors else None, _vec_dtype(x), _stream_handle(stream))
        N.check(rc, "hmvm_grad")
        return garena


def _check_buffer(t, device: int, nbytes: int, what: str) -> None:
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous torch tensor")
    if t.numel() * t.element_size() < nbytes:
        raise ValueError(f"{what}: {t.numel() * t.element_size()} bytes, {nbytes} needed")
    _on_device(t, device, what)


def _f64_vector(t, n: int, what: str):
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or t.dim() != 1 or t.numel() != n \
            or not t.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous float64 device vector of {n} entries")
    return C.c_void_p(t.data_ptr())


def inv_error_2(A: Plan, W: Plan, atol: float = 0.0, max_it: int = 0, stream=None):
    """|I - A W|_2 by power iteration on the device (the reference's
    hlr::norm::inv_error_2, arith/norm.hh:115-137, which runs norm::spectral
    on I - A A_inv): returns (error, iterations).  atol = 0: sqrt(eps);
    max_it = 0: max(5, n / 10)."""
    err = C.c_double()
    its = C.c_int64()
    N.check(N.hmvm().hmvm_inv_error_2(A.handle, W.handle, float(atol), int(max_it), _stream_handle(stream),
                                      C.byref(err), C.byref(its)), "hmvm_inv_error_2")
    return float(err.value), int(its.value)


def linear_iteration(A: Plan, W: Plan, b, x, sol=None, tol: float = 0.0, max_steps: int = 0, stream=None):
    """x <- x - W (A x - b) on the device until |sol - x| <= tol (sol given)
    or |A x - b| <= tol |b|, or max_steps steps (the reference's
    linear_iteration, programs/iterrefine.hh:20-80; tol = 0: 1e-8,
    max_steps = 0: 100).  b, x, sol: float64 tensors on the plans' device;
    x holds the start value and receives the result.  Returns (steps,
    history), history[i] = (|r| before step i, |c|, |e| after it)."""
    n = A.nrows
    limit = int(max_steps) if max_steps else 100
    hist = np.full(3 * max(limit, 1), np.nan)
    steps = C.c_int64()
    N.check(N.hmvm().hmvm_linear_iteration(A.handle, W.handle, _f64_vector(b, n, "b"), _f64_vector(x, n, "x"),
                                           _f64_vector(sol, n, "sol") if sol is not None else None, float(tol),
                                           int(max_steps), _stream_handle(stream), C.byref(steps),
                                           hist.ctypes.data_as(C.POINTER(C.c_double))), "hmvm_linear_iteration")
    return int(steps.value), hist[:3 * steps.value].reshape(-1, 3)
