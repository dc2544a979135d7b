"""Batched products (GPU): many Matern H-matrices of uneven size (so panels
with fewer than 64 rows occur) and right-hand sides in {1,4,16,64}, run as
one stream-ordered launch sequence; checked member by member against the
oracle, recorded in a HIP graph and replayed several times with eager work
in between, and run again after the autotuner picked a schedule.  The
per-plan product on the same members runs first, so that a failure of the
batched launches points at the batch indexing, not at the shared routines."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.batch import Batch
from hlr_amd.hmatrix import HMatrix, batch_problems, random_vector
from hlr_amd.plan import Plan
from tests.helpers import rel

pytestmark = pytest.mark.gpu

NRHS = [1, 4, 16, 64, 1, 4, 16, 64]


@pytest.fixture(scope="module")
def members():
    probs = batch_problems(count=8, seed=3, nmin=1500, nmax=6000)
    hms = [HMatrix(p) for p, _ in probs]
    return hms, list(NRHS)


def _vectors(hms, nrhs, cuda, seed):
    import torch
    xs, ys = [], []
    for i, (hm, r) in enumerate(zip(hms, nrhs)):
        X = np.stack([random_vector(hm.n, seed + 100 * i + j, gaussian=True) for j in range(r)])
        xs.append(torch.as_tensor(X if r > 1 else X[0], device=cuda))
        ys.append(torch.zeros_like(xs[-1]))
    return xs, ys


def _check(hms, nrhs, xs, ys, alphas, tol):
    for hm, r, x, y, a in zip(hms, nrhs, xs, ys, alphas):
        X = x.cpu().numpy()
        ref = np.zeros_like(X)
        oracle.mul_vec(hm.blocks, hm.arena, a, N.OP_NORMAL, X, ref)
        assert rel(y.cpu().numpy(), ref) <= tol, (hm.n, r)


def test_members_per_plan(cuda, members):
    hms, nrhs = members
    xs, ys = _vectors(hms, nrhs, cuda, 5)
    alphas = [0.5 + i for i in range(len(hms))]
    for hm, r, x, y, a in zip(hms, nrhs, xs, ys, alphas):
        print(f"[gpu] per-plan n={hm.n} nrhs={r}", flush=True)
        p = Plan.from_hmatrix(hm, precision="fp64")
        p.apply(a, N.OP_NORMAL, x, y)
    _check(hms, nrhs, xs, ys, alphas, 1e-13)


def test_batch_apply_graph_tune(cuda, members):
    import torch
    hms, nrhs = members
    plans = [Plan.from_hmatrix(hm, precision="fp64") for hm in hms]
    b = Batch(plans, nrhs)
    xs, ys = _vectors(hms, nrhs, cuda, 7)
    alphas = [1.0 + 0.5 * i for i in range(len(hms))]
    print(f"[gpu] batch n={[hm.n for hm in hms]} nrhs={nrhs}", flush=True)
    b.apply(xs, ys, alphas)
    b.check()
    _check(hms, nrhs, xs, ys, alphas, 1e-13)

    # graph capture and replays; eager work and synchronisation in between
    print("[gpu] batch graph capture", flush=True)
    for y in ys:
        y.zero_()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        b.capture(xs, ys, alphas, stream=stream)
    stream.synchronize()
    for rep in range(3):
        for y in ys:
            y.zero_()
        torch.cuda.synchronize()
        tmp = torch.randn(1 << 20, device=cuda).sum()      # unrelated eager work
        b.replay(stream=stream)
        b.check(stream=stream)
        _check(hms, nrhs, xs, ys, alphas, 1e-13)
        del tmp

    # autotuner, then a plain launch with the chosen schedule
    print("[gpu] batch tune", flush=True)
    chosen = b.tune(xs, ys)
    assert 0 <= chosen < 4
    for y in ys:
        y.zero_()
    b.apply(xs, ys, alphas)
    b.check()
    _check(hms, nrhs, xs, ys, alphas, 1e-13)
    b.close()


def test_batch_mixed_precision(cuda, members):
    hms, nrhs = members
    plans = [Plan.from_hmatrix(hm, precision="valr", eps=1e-6) for hm in hms]
    b = Batch(plans, nrhs)
    xs, ys = _vectors(hms, nrhs, cuda, 9)
    print(f"[gpu] batch valr n={[hm.n for hm in hms]} nrhs={nrhs}", flush=True)
    b.apply(xs, ys)
    b.check()
    _check(hms, nrhs, xs, ys, [1.0] * len(hms), 1e-6)
    b.close()


@pytest.mark.parametrize("prec", ["bfp", "afl"])
def test_batch_packed_formats(cuda, members, prec):
    """Members stored in block floating point or AFL: the batch runs the
    family's own kernel instantiation and stays within eps of the oracle."""
    hms, nrhs = members
    plans = [Plan.from_hmatrix(hm, precision=prec, eps=1e-6) for hm in hms]
    b = Batch(plans, nrhs)
    xs, ys = _vectors(hms, nrhs, cuda, 10)
    print(f"[gpu] batch {prec} n={[hm.n for hm in hms]} nrhs={nrhs}", flush=True)
    b.apply(xs, ys)
    b.check()
    _check(hms, nrhs, xs, ys, [1.0] * len(hms), 1e-6)
    b.close()


def test_batch_graph_keeps_its_vectors(cuda, members):
    """A captured graph replays the product it was captured with: after an
    eager apply with other vectors and alphas, and after the autotuner, a
    replay writes the captured y (checked against the oracle) and leaves the
    other call's y untouched."""
    import torch
    hms, nrhs = members
    plans = [Plan.from_hmatrix(hm, precision="fp64") for hm in hms]
    b = Batch(plans, nrhs)
    xs1, ys1 = _vectors(hms, nrhs, cuda, 21)
    xs2, ys2 = _vectors(hms, nrhs, cuda, 22)
    a1 = [0.25 + i for i in range(len(hms))]
    a2 = [-3.0 - i for i in range(len(hms))]
    print("[gpu] batch capture, then apply and tune with other vectors, then replay", flush=True)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        b.capture(xs1, ys1, a1, stream=stream)
    stream.synchronize()
    b.apply(xs2, ys2, a2)
    b.check()
    _check(hms, nrhs, xs2, ys2, a2, 1e-13)
    scratch = [torch.zeros_like(y) for y in ys2]
    b.tune(xs2, scratch)
    torch.cuda.synchronize()
    before = [y.clone() for y in ys2]
    for y in ys1:
        y.zero_()
    torch.cuda.synchronize()
    b.replay(stream=stream)
    b.check(stream=stream)
    _check(hms, nrhs, xs1, ys1, a1, 1e-13)
    for y, y0 in zip(ys2, before):
        assert torch.equal(y, y0)
    b.close()
