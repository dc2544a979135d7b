"""The Python bindings reject bad vectors before any C call.

The C-ABI takes raw device pointers and cannot check lengths, dtypes or
strides itself, so Plan.apply / apply_ws / apply_timed / tlr_apply / grad /
set_arena and Batch check them.  These tests run on the CPU: a host-only plan
(device = -1) rejects every CPU tensor on its device check, which comes after
the shape, dtype and stride checks, so each bad case is told apart by its
message."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from hlr_amd import _native as N
from hlr_amd.batch import Batch
from hlr_amd.hmatrix import HMatrix, LOG1D
from hlr_amd.plan import Plan, check_vectors


@pytest.fixture(scope="module")
def plan():
    hm = HMatrix(LOG1D.with_(n=512))
    p = Plan.from_hmatrix(hm, precision="fp64", device=-1)
    yield p
    p.close()
    hm.close()


def _v(n, dtype=torch.float64):
    return torch.zeros(n, dtype=dtype)


def _m(r, n):
    return torch.zeros(r, n, dtype=torch.float64)


def test_well_formed_vectors_reach_the_device_check(plan):
    n = plan.nrows
    with pytest.raises(ValueError, match="HIP device -1"):
        check_vectors(plan, N.OP_NORMAL, _v(n), _v(n))
    with pytest.raises(ValueError, match="HIP device -1"):
        check_vectors(plan, N.OP_NORMAL, _m(3, n), _m(3, n))


@pytest.mark.parametrize("case,exc,msg", [
    ("short_x", ValueError, "expected x of"),
    ("long_y", ValueError, "expected x of"),
    ("dtype_mix", ValueError, "but y is"),
    ("int_dtype", TypeError, "unsupported vector dtype"),
    ("strided", ValueError, "unit stride"),
    ("dims", ValueError, "both be 1-D"),
    ("rhs_mismatch", ValueError, "same number"),
    ("rows_short", ValueError, r"expected \(nrhs"),
    ("not_tensor", TypeError, "not a torch tensor"),
])
def test_rejected(plan, case, exc, msg):
    n = plan.nrows
    x, y = _v(n), _v(n)
    if case == "short_x":
        x = _v(n - 1)
    elif case == "long_y":
        y = _v(n + 1)
    elif case == "dtype_mix":
        y = _v(n, torch.float32)
    elif case == "int_dtype":
        x, y = torch.zeros(n, dtype=torch.int64), torch.zeros(n, dtype=torch.int64)
    elif case == "strided":
        x = torch.zeros(2 * n, dtype=torch.float64)[::2]
    elif case == "dims":
        y = torch.zeros(1, n, dtype=torch.float64)
    elif case == "rhs_mismatch":
        x, y = torch.zeros(3, n, dtype=torch.float64), torch.zeros(2, n, dtype=torch.float64)
    elif case == "rows_short":
        x, y = torch.zeros(2, n - 1, dtype=torch.float64), torch.zeros(2, n, dtype=torch.float64)
    elif case == "not_tensor":
        x = np.zeros(n)
    with pytest.raises(exc, match=msg):
        check_vectors(plan, N.OP_NORMAL, x, y)
    # every entry point that takes vectors goes through the same check
    with pytest.raises(exc, match=msg):
        plan.apply(1.0, N.OP_NORMAL, x, y)


def test_expected_rhs_count(plan):
    n = plan.nrows
    with pytest.raises(ValueError, match="expected 1 right-hand sides"):
        check_vectors(plan, N.OP_NORMAL, _m(2, n), _m(2, n), nrhs=1)
    with pytest.raises(ValueError, match="one right-hand side"):
        plan.tlr_apply(1.0, N.OP_NORMAL, _m(1, n), _m(1, n))


def test_workspace_and_arena_buffers(plan):
    n = plan.nrows
    with pytest.raises(ValueError, match="expected x of"):
        plan.apply_ws(1.0, N.OP_NORMAL, _v(n - 2), _v(n), torch.zeros(8))
    with pytest.raises(ValueError, match="bytes"):
        plan.set_arena(N.FMT_F64, torch.zeros(1))


def test_grad_vectors(plan):
    n = plan.nrows
    with pytest.raises(ValueError, match="1-D"):
        plan.grad(_m(1, n), _v(n))
    with pytest.raises(ValueError, match="but y is"):
        plan.grad(_v(n), _v(n, torch.float32))
    with pytest.raises(ValueError, match="gx must be"):
        plan.grad(_v(n), _v(n), gx=_v(n - 1))


def _bare_batch(plans, nrhs, vdt=N.F64):
    # the checks of Batch._args without a device batch behind them
    b = object.__new__(Batch)
    b._b = None
    b.plans, b.nrhs, b.vec_dtype = list(plans), list(nrhs), vdt
    return b


def test_batch_member_counts(plan):
    n = plan.nrows
    b = _bare_batch([plan, plan], [1, 3])
    with pytest.raises(ValueError, match="2 members, got 1 x"):
        b._args(None, [_v(n)], [_v(n), _v(n)])
    with pytest.raises(ValueError, match="got 3 alphas"):
        b._args([1.0, 2.0, 3.0], [_v(n), _m(3, n)], [_v(n), _m(3, n)])
    with pytest.raises(ValueError, match="expected 3 right-hand sides"):
        b._args(None, [_v(n), _m(2, n)], [_v(n), _m(2, n)])
    with pytest.raises(ValueError, match="member 0: expected x of"):
        b._args(None, [_v(n - 1), _m(3, n)], [_v(n), _m(3, n)])
    b32 = _bare_batch([plan], [1], N.F32)
    with pytest.raises(ValueError, match="HIP device -1|created for"):
        b32._args(None, [_v(n)], [_v(n)])
