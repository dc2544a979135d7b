"""Batched persistent launch (GPU): many Matern H-matrices of uneven size
(so panels with fewer than 64 rows occur), right-hand sides in {1,4,16,64},
one launch; checked member by member against the oracle, then recorded in a
HIP graph and replayed several times with eager work in between, and run
again after the autotuner picked a launch shape."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hlr_amd import _native as N
from hlr_amd.batch import Batch
from hlr_amd.hmatrix import HMatrix, batch_problems, random_vector
from hlr_amd.plan import Plan
from tests.helpers import rel

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="written for the single persistent-kernel batch launch, whose kernel did not "
                            "finish on the MI355X (twice, also with every inter-workgroup wait bounded); "
                            "no cause was found by reading the code, so that kernel was taken out and the "
                            "batch API now runs stream-ordered launches, tested in "
                            "test_gpu_batch_launches.py"),
]


@pytest.fixture(scope="module")
def members():
    probs = batch_problems(count=8, seed=3, nmin=1500, nmax=6000)
    # make sure every right-hand-side class occurs
    probs = [(p, r) for (p, _), r in zip(probs, [1, 4, 16, 64, 1, 4, 16, 64])]
    hms = [HMatrix(p) for p, _ in probs]
    return hms, [r for _, r in probs]


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
        stream.synchronize()
        b.check(stream=stream)
        _check(hms, nrhs, xs, ys, alphas, 1e-13)
        del tmp

    # autotuner, then a plain launch with the chosen shape
    print("[gpu] batch tune", flush=True)
    chosen = b.tune(xs, ys)
    assert 0 <= chosen < 8
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
