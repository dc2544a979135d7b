#!/usr/bin/env python3
"""H-matrix-vector product throughput on one MI355X.

Workload: kernel 1/|x-y| on n = 262144 points uniform on the unit sphere
(diagonal = mean of the 8 nearest-neighbour entries), binary space
partitioning with leaf size 64, standard admissibility eta = 2, ACA + SVD
truncation to eps = 1e-6, x uniform in [-1, 1], nrhs = 1.

One step = one H-MVM y += 2 A x over the whole matrix.  The mixed-precision
plan (per-column FP32/BF16/FP8-E4M3 chosen from singular values against eps,
dense leaves FP32, FP32 accumulation inside panels) is the measured path;
the plain FP64 plan is timed beside it, both with factors resident in HBM.
A third plan stores every column in block floating point (values of 8, 16 or 25
bits with a power-of-two step per panel column, width chosen per column so
that the column's 2-norm error stays within the same VALR budget); its
figures are reported beside the mixed-precision ones under "bfp".  A fourth
stores them in adaptive float codes of 12, 16 or 24 bits (sign, 5 exponent
bits, truncated mantissa, an FP32 scale per panel column) under "afl".
The relative error of the mixed-precision result is taken against the FP64
GPU result (which the full-size parity tests hold to <= 1e-13 of the FP64
oracle on this matrix).  The CPU baseline is the oracle's restatement of the
reference's TBB mul_vec_chunk on the host cores.  --nrhs R > 1 times R
right-hand sides per call (matrix-core kernels); one step is then one call
and value counts R products per call.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed H-MVMs per precision")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nrhs", type=int, default=1, help="right-hand sides per call")
    ap.add_argument("--op", choices=["normal", "transposed"], default="normal",
                    help="product timed (the square sphere matrix keeps the same sizes)")
    ap.add_argument("--precision", choices=["both", "fp64", "valr", "bfp", "afl"], default="both",
                    help="time only one plan (profiling runs); the JSON line needs fp64 and valr")
    ap.add_argument("--no-bfp", action="store_true", help="skip the block floating point plan")
    ap.add_argument("--no-afl", action="store_true", help="skip the adaptive float plan")
    ap.add_argument("--mrhs", type=int, default=32,
                    help="also time each plan with this many right-hand sides per call (0: skip)")
    ap.add_argument("--copy-gib", type=float, default=2.0, help="size of the device copy calibration")
    ap.add_argument("--workload", choices=["sphere", "matern", "batch", "grad", "uniform", "construct"],
                    default="sphere",
                    help="sphere: 1/r, n = 262144 (default); matern: the plain FP64 drop-in on n = 65536; "
                         "batch: 64 Matern matrices, many right-hand sides; grad: loss and gradients (TLR); "
                         "uniform: uniform-H and H^2 forms of the sphere matrix at n = 65536; "
                         "construct: building the sphere matrix with its low-rank leaves on the GPU")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the H-MVM path here is single-GPU "
              "(sharding over several GPUs is not built).", file=sys.stderr)
        return 2
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "matern":
        return bench_matern(args)
    if args.workload == "uniform":
        return bench_uniform(args)
    if args.workload == "grad":
        return bench_grad(args)
    if args.workload == "construct":
        return bench_construct(args)

    import numpy as np
    import torch

    from hlr_amd import _native as N
    from hlr_amd.hmatrix import INVR_SPHERE, HMatrix, random_vector
    from hlr_amd.plan import Plan

    dev = torch.device("cuda:0")
    prob = INVR_SPHERE.with_(n=args.n, eps=args.eps)
    t0 = time.time()
    hm = HMatrix(prob)
    st = hm.stats()
    log(f"matrix n={hm.n} leaves={st['leaves']} dense={st['dense']} lowrank={st['lowrank']} "
        f"kmax={st['kmax']} fp64 bytes={(st['dense_bytes_fp64'] + st['lowrank_bytes_fp64']) / 1e9:.2f} GB "
        f"built in {time.time() - t0:.1f} s")

    R = args.nrhs
    OP = N.OP_TRANSPOSED if args.op == "transposed" else N.OP_NORMAL
    if OP != N.OP_NORMAL and R > 1:
        log("transposed products take one right-hand side per call here")
        return 2
    x_h = random_vector(hm.n, 2) if R == 1 else np.stack([random_vector(hm.n, 2 + j) for j in range(R)])
    x = torch.as_tensor(x_h, device=dev)
    stream = torch.cuda.current_stream()

    def time_plan(plan, steps, warmup):
        # workspace allocated once (hmvm_apply_ws): the timed loop is kernels only
        ws = torch.empty(plan.workspace_bytes(OP, R) // 8 + 1, dtype=torch.float64, device=dev)
        y = torch.zeros_like(x)
        for _ in range(warmup):
            plan.apply_ws(2.0, OP, x, y, ws)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(steps):
            plan.apply_ws(2.0, OP, x, y, ws)
        stop.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / steps
        if R > 1:
            return ms, np.zeros(3)
        # per-phase durations (HIP events on the launch stream)
        phases = np.array([plan.apply_timed(2.0, OP, x, y) for _ in range(max(5, steps // 10))])
        return ms, phases.mean(axis=0)

    def result(plan):
        y = torch.zeros_like(x)
        plan.apply(1.0, OP, x, y)
        torch.cuda.synchronize()
        return y.cpu().numpy()

    # several right-hand sides per call (matrix cores), beside the single-RHS
    # line: ms per call and the product, for each plan before it is closed
    RM = args.mrhs if (R == 1 and OP == N.OP_NORMAL) else 0
    xm = torch.as_tensor(np.stack([random_vector(hm.n, 500 + j) for j in range(RM)]), device=dev) if RM > 1 else None

    def time_mrhs(plan, steps=10, warmup=2):
        if xm is None:
            return None
        ws = torch.empty(plan.workspace_bytes(N.OP_NORMAL, RM) // 8 + 1, dtype=torch.float64, device=dev)
        y = torch.zeros_like(xm)
        for _ in range(warmup):
            plan.apply_ws(2.0, N.OP_NORMAL, xm, y, ws)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(steps):
            plan.apply_ws(2.0, N.OP_NORMAL, xm, y, ws)
        stop.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / steps
        y.zero_()
        plan.apply(1.0, N.OP_NORMAL, xm, y)
        torch.cuda.synchronize()
        out = (ms, y.cpu().numpy())
        del ws, y
        torch.cuda.empty_cache()
        return out

    # ---- device copy of known size: achievable HBM bandwidth on this GPU ----
    nbytes = int(args.copy_gib * 2 ** 30)
    src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    c0 = torch.cuda.Event(enable_timing=True)
    c1 = torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    for _ in range(10):
        dst.copy_(src)
    c1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 2.0 * nbytes / (c0.elapsed_time(c1) / 10 * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    log(f"device copy: {copy_gbs:.0f} GB/s (read + write of {args.copy_gib:g} GiB)")

    ms64 = msmp = None
    m64 = mmp = None
    ph64 = phmp = np.zeros(3)
    i64 = imp = None
    y64 = ymp = None
    # ---- plain FP64 plan ----
    if args.precision in ("both", "fp64"):
        t0 = time.time()
        p64 = Plan.from_hmatrix(hm, precision="fp64")
        i64 = p64.info()
        log(f"fp64 plan: {i64['factor_bytes'] / 1e9:.2f} GB in {time.time() - t0:.1f} s")
        ms64, ph64 = time_plan(p64, args.steps, args.warmup)
        y64 = result(p64)
        log(f"fp64: {ms64:.3f} ms/MVM  phases {ph64.round(4).tolist()} ms")
        m64 = time_mrhs(p64)
        p64.close()
        torch.cuda.empty_cache()

    # ---- mixed-precision plan ----
    if args.precision in ("both", "valr"):
        t0 = time.time()
        pmp = Plan.from_hmatrix(hm, precision="valr", eps=args.eps)
        imp = pmp.info()
        log(f"valr plan: {imp['factor_bytes'] / 1e9:.2f} GB cols/format {imp['cols_by_format']} "
            f"in {time.time() - t0:.1f} s")
        msmp, phmp = time_plan(pmp, args.steps, args.warmup)
        ymp = result(pmp)
        log(f"valr: {msmp:.3f} ms/MVM  phases {phmp.round(4).tolist()} ms")
        mmp = time_mrhs(pmp)
        pmp.close()

    # ---- packed plans: block floating point, adaptive float ----
    packed = {}
    for fam, skip in (("bfp", args.no_bfp), ("afl", args.no_afl)):
        if not (args.precision == fam or (args.precision == "both" and not skip)):
            continue
        t0 = time.time()
        pk = Plan.from_hmatrix(hm, precision=fam, eps=args.eps)
        ipk = pk.info()
        log(f"{fam} plan: {ipk['factor_bytes'] / 1e9:.2f} GB cols/format {ipk['cols_by_format']} "
            f"in {time.time() - t0:.1f} s")
        mspk, phpk = time_plan(pk, args.steps, args.warmup)
        ypk = result(pk)
        log(f"{fam}: {mspk:.3f} ms/MVM  phases {phpk.round(4).tolist()} ms")
        mpk = time_mrhs(pk)
        pk.close()
        torch.cuda.empty_cache()
        packed[fam] = {"ms_per_step": mspk, "info": ipk, "phases": phpk, "y": ypk, "mrhs": mpk}

    if args.precision != "both":
        log("single-precision profiling run: no JSON line")
        return 0

    # relative error per right-hand side, worst over the right-hand sides
    d = (ymp - y64).reshape(R, -1)
    err_gpu = float(max(np.linalg.norm(d[j]) / np.linalg.norm(y64.reshape(R, -1)[j]) for j in range(R)))
    cpu = None
    if not args.no_cpu and R == 1 and OP == N.OP_NORMAL:
        cpu = cpu_baseline(hm.blocks, hm.arena, x_h, hm.n, args.cpu_seconds)
    # ---- roofline of the dominant kernel: the row-panel pass (k_ufwd) ----
    n = hm.n
    mvm_bytes = imp["datasize"] + 8.0 * 2 * (R - 1) * (n + n)    # vectors counted per right-hand side
    if R == 1 and OP == N.OP_NORMAL:
        vec_bytes = 8.0 * (n + n)               # read x, read+write y once per pass
        r_bytes = imp["bytes_rpass"] + vec_bytes
        achieved = r_bytes / (phmp[2] * 1e-3) / 1e9
        rkernel = "k_ufwd (row-panel pass, mixed precision)"
    elif R == 1:
        # transposed: the row-panel pass is the first phase and reads x once
        r_bytes = imp["bytes_rpass"] + 8.0 * n
        achieved = r_bytes / (phmp[0] * 1e-3) / 1e9
        rkernel = "k_rtrn (row-panel pass of the transposed product, mixed precision)"
    else:
        # whole call: the matrix-core kernels stream the factors once per call
        r_bytes = mvm_bytes
        achieved = r_bytes / (msmp * 1e-3) / 1e9
        rkernel = f"whole call ({R} right-hand sides, matrix-core kernels)"
    gbs_mp = mvm_bytes / (msmp * 1e-3) / 1e9
    gbs_64 = (i64["datasize"] + 8.0 * 2 * (R - 1) * (n + n)) / (ms64 * 1e-3) / 1e9

    # HBM traffic of the same kernel from the committed rocprofv3 PMC summary
    # (FETCH_SIZE + WRITE_SIZE per launch, copy-calibrated), when it exists
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_invr_sphere_n{n}.txt"
    if tfile.exists() and args.eps == 1e-6 and R == 1 and OP == N.OP_NORMAL:
        for line in tfile.read_text().splitlines():
            f = line.split()
            if len(f) >= 4 and not line.startswith("#") and f[0] == "valr" and f[1].startswith("k_ufwd"):
                traffic = float(f[-2]) + float(f[-1])

    out = {
        "metric": "H-MVM/s + HBM GB/s (% peak), mixed-prec vs FP64 GPU at rel.err<=eps, n=262k",
        "value": 1000.0 * R / msmp,
        "unit": "H-MVM/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": msmp,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-acc mixed fp32/bf16/fp8-e4m3 storage",
        "data": "synthetic (seeded points, generated H-matrix)",
        "config": {"workload": f"1/r on unit sphere, n={n}, leaf 64, eta 2, ACA eps={args.eps:g}, nrhs {R}"
                               + ("" if OP == N.OP_NORMAL else ", transposed product"),
                   "precision": "valr", "parallelism": "single GPU"},
        "fp64_mvm_per_s": 1000.0 * R / ms64,
        "fp64_ms_per_step": ms64,
        "speedup_vs_fp64": ms64 / msmp,
        "rel_err_vs_fp64_gpu": err_gpu,
        "hbm_gbs_mixed": gbs_mp,
        "hbm_gbs_fp64": gbs_64,
        "hbm_frac_mixed": gbs_mp / HBM_PEAK_GBS,
        "hbm_frac_fp64": gbs_64 / HBM_PEAK_GBS,
        "device_copy_gbs": copy_gbs,
        "bytes_per_mvm_mixed": mvm_bytes,
        "bytes_per_mvm_fp64": i64["datasize"],
        "phases_ms_mixed": phmp.tolist(),
        "phases_ms_fp64": ph64.tolist(),
        "roofline": {"bound": "hbm", "kernel": rkernel,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": (f"profiles/{tfile.name} (rocprofv3 PMC FETCH_SIZE + WRITE_SIZE of the "
                                        "same kernel and workload, copy-calibrated; not measured in this run)")
                     if traffic is not None else None,
                     "algorithmic_bytes": r_bytes},
        "cpu_baseline": cpu,
    }
    for fam, pk in packed.items():
        db = (pk["y"] - y64).reshape(R, -1)
        err_pk = float(max(np.linalg.norm(db[j]) / np.linalg.norm(y64.reshape(R, -1)[j]) for j in range(R)))
        ib = pk["info"]
        bb = ib["datasize"] + 8.0 * 2 * (R - 1) * (n + n)
        bf = {"mvm_per_s": 1000.0 * R / pk["ms_per_step"], "ms_per_step": pk["ms_per_step"],
              "speedup_vs_fp64": ms64 / pk["ms_per_step"], "rel_err_vs_fp64_gpu": err_pk,
              "bytes_per_mvm": bb, "hbm_gbs": bb / (pk["ms_per_step"] * 1e-3) / 1e9,
              "phases_ms": pk["phases"].tolist(), "cols_by_format": ib["cols_by_format"]}
        if R == 1 and OP == N.OP_NORMAL and pk["phases"][2] > 0:
            rb = ib["bytes_rpass"] + 8.0 * (n + n)
            bf["rpass_gbs"] = rb / (pk["phases"][2] * 1e-3) / 1e9
        out[fam] = bf
    if RM > 1 and m64 is not None and mmp is not None:
        # worst relative error over the right-hand sides against the FP64 plan
        def worst(m):
            return float(max(np.linalg.norm(m[1][j] - m64[1][j]) / np.linalg.norm(m64[1][j]) for j in range(RM)))
        mr = {"nrhs": RM, "fp64_ms": m64[0], "valr_ms": mmp[0], "speedup_valr": m64[0] / mmp[0],
              "rel_err_valr": worst(mmp), "h_mvm_per_s_fp64": 1000.0 * RM / m64[0],
              "h_mvm_per_s_valr": 1000.0 * RM / mmp[0]}
        for fam, pk in packed.items():
            mpk = pk["mrhs"]
            if mpk is not None:
                mr.update({f"{fam}_ms": mpk[0], f"speedup_{fam}": m64[0] / mpk[0], f"rel_err_{fam}": worst(mpk),
                           f"h_mvm_per_s_{fam}": 1000.0 * RM / mpk[0]})
        out["mrhs"] = mr
        log(f"{RM} rhs: fp64 {m64[0]:.3f} ms, valr {mmp[0]:.3f} ms"
            + "".join(f", {fam} {mr[fam + '_ms']:.3f} ms" for fam in packed if f"{fam}_ms" in mr))
    print(json.dumps(out), flush=True)
    return 0


def bench_matern(args) -> int:
    """Plain FP64 drop-in: Matern nu = 1.5, l = 0.1, unit variance on
    n = 65536 points in the unit cube, leaf 64, eta 2, ACA eps 1e-6, FP64
    factors and accumulation, nrhs 1, x ~ U[-1, 1].  One step is one
    y += 2 A x with the factors resident in HBM; the relative error is
    taken against the oracle's FP64 mul_vec on the same matrix, and the
    restatement of tbb::mul_vec_chunk is timed on the host cores beside it."""
    import numpy as np
    import torch

    import oracle
    from hlr_amd import _native as N
    from hlr_amd.hmatrix import MATERN_CUBE, HMatrix, random_vector
    from hlr_amd.plan import Plan

    dev = torch.device("cuda:0")
    t0 = time.time()
    hm = HMatrix(MATERN_CUBE)
    st = hm.stats()
    log(f"matrix n={hm.n} leaves={st['leaves']} dense={st['dense']} lowrank={st['lowrank']} kmax={st['kmax']} "
        f"built in {time.time() - t0:.1f} s")
    x_h = random_vector(hm.n, 2)
    x = torch.as_tensor(x_h, device=dev)
    plan = Plan.from_hmatrix(hm, precision="fp64")
    info = plan.info()
    ws = torch.empty(plan.workspace_bytes(N.OP_NORMAL, 1) // 8 + 1, dtype=torch.float64, device=dev)
    y = torch.zeros_like(x)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        plan.apply_ws(2.0, N.OP_NORMAL, x, y, ws)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        plan.apply_ws(2.0, N.OP_NORMAL, x, y, ws)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    phases = np.array([plan.apply_timed(2.0, N.OP_NORMAL, x, y) for _ in range(max(5, args.steps // 10))]).mean(axis=0)
    # error against the oracle (the checker) on the same matrix
    y1 = torch.zeros_like(x)
    plan.apply(1.0, N.OP_NORMAL, x, y1)
    torch.cuda.synchronize()
    yr = np.zeros(hm.n)
    oracle.mul_vec(hm.blocks, hm.arena, 1.0, N.OP_NORMAL, x_h, yr)
    err = float(np.linalg.norm(y1.cpu().numpy() - yr) / np.linalg.norm(yr))
    cpu = None if args.no_cpu else cpu_baseline(hm.blocks, hm.arena, x_h, hm.n, args.cpu_seconds)
    n = hm.n
    r_bytes = info["bytes_rpass"] + 8.0 * (n + n)
    achieved = r_bytes / (phases[2] * 1e-3) / 1e9
    gbs = info["datasize"] / (ms * 1e-3) / 1e9
    log(f"matern fp64: {ms:.4f} ms/MVM, phases {phases.round(4).tolist()} ms, {gbs:.0f} GB/s, rel.err {err:.2e}")
    out = {
        "metric": "H-MVM/s, plain FP64 drop-in (Matern n=65536) vs the CPU restatement",
        "value": 1000.0 / ms, "unit": "H-MVM/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded points, generated H-matrix)",
        "config": {"workload": f"Matern nu=1.5 l=0.1 on n={n} points in the unit cube, leaf 64, eta 2, "
                               "ACA eps 1e-6, FP64, nrhs 1", "parallelism": "single GPU"},
        "rel_err_vs_oracle": err, "bytes_per_mvm": info["datasize"], "hbm_gbs": gbs,
        "hbm_frac": gbs / HBM_PEAK_GBS, "phases_ms": phases.tolist(),
        "roofline": {"bound": "hbm", "kernel": "k_ufwd (row-panel pass, FP64)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": r_bytes},
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        out["speedup_vs_cpu"] = out["value"] / cpu["value"]
        out["speedup_vs_cpu_1core"] = out["value"] / cpu["value_1core"]
    print(json.dumps(out), flush=True)
    return 0


def bench_construct(args) -> int:
    """Construction of the 1/r sphere matrix (n = 262144, leaf 64, eta 2, ACA
    eps 1e-6): the host generator on all host cores (tree, blocks, dense
    leaves and ACA + SVD truncation of the low-rank leaves) against the
    split build (host: tree, blocks and dense leaves; GPU: ACA + truncation
    of the low-rank leaves, hmvm_lowrank_build).  One step = one whole
    construction; the split build's product is checked against the host
    build's."""
    import numpy as np
    import torch

    from hlr_amd import _native as N
    from hlr_amd.hmatrix import INVR_SPHERE, HMatrix, random_vector
    from hlr_amd.plan import Plan

    dev = torch.device("cuda:0")
    torch.zeros(1, device=dev)                      # runtime up before the clocks
    prob = INVR_SPHERE.with_(n=args.n)
    warm = min(args.warmup, 2)                      # whole constructions: a few suffice
    for _ in range(warm):
        HMatrix(prob.with_(n=8192), construct="device").close()
    t0 = time.time()
    hm = HMatrix(prob)
    t_host = time.time() - t0
    walls, devs, kers, cops = [], [], [], []
    hd = None
    for _ in range(max(1, min(args.steps, 3))):
        if hd is not None:
            hd.close()
        t0 = time.time()
        hd = HMatrix(prob, construct="device")
        walls.append(time.time() - t0)
        devs.append(hd.device_seconds)
        kers.append(hd.device_kernel_seconds)
        cops.append(hd.device_copy_seconds)
    lr = hm.blocks["kind"] != N.BLOCK_DENSE
    same = float((hd.blocks["rank"][lr] == hm.blocks["rank"][lr]).mean())
    x = torch.as_tensor(random_vector(hm.n, 3), device=dev)
    ys = []
    for h in (hm, hd):
        p = Plan.from_hmatrix(h, precision="fp64")
        y = torch.zeros_like(x)
        p.apply(1.0, N.OP_NORMAL, x, y)
        torch.cuda.synchronize()
        ys.append(y.cpu().numpy())
        p.close()
    err = float(np.linalg.norm(ys[1] - ys[0]) / np.linalg.norm(ys[0]))
    wall = float(np.median(walls))
    lr_bytes = float(sum(8.0 * (int(b["nrows"]) + int(b["ncols"]) + 1) * int(b["rank"]) for b in hd.blocks[lr]))
    ncpu = len(os.sched_getaffinity(0))
    log(f"construct: host {t_host:.2f} s ({hm.build_seconds:.2f} s in hlrgen_build, {ncpu} cores), "
        f"split {wall:.2f} s (GPU part {np.median(devs):.3f} s: kernels {np.median(kers):.3f} s, factors to the "
        f"host {np.median(cops):.3f} s), ranks equal on {same:.4f} of the low-rank leaves, "
        f"product difference {err:.2e}")
    out = {
        "metric": "H-matrix construction time (1/r sphere, n=262144), low-rank leaves on the GPU",
        "value": wall, "unit": "s", "n_gpus": 1, "steps": len(walls), "warmup": warm, "ms_per_step": wall * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded points on the sphere)",
        "config": {"workload": f"1/r on unit sphere, n={hm.n}, leaf 64, eta 2, ACA eps 1e-6: tree, blocks and dense "
                               "leaves on the host, ACA + SVD truncation of the low-rank leaves on the GPU",
                   "parallelism": "single GPU"},
        "gpu_seconds_lowrank": float(np.median(devs)),
        "gpu_kernel_seconds": float(np.median(kers)), "factor_copy_seconds": float(np.median(cops)),
        "lowrank_leaves": int(lr.sum()), "lowrank_factor_bytes": lr_bytes,
        "ranks_equal_to_host_build": same, "product_rel_diff_vs_host_build": err,
        "roofline": None,
        "cpu_baseline": {"value": t_host, "unit": "s", "cores": ncpu, "kind": "port",
                         "sample": f"the host generator (hlrgen_build, C++/OpenMP, all {ncpu} cores) building the "
                                   "same matrix, ACA and truncation included (one build)"},
    }
    print(json.dumps(out), flush=True)
    hd.close()
    hm.close()
    return 0


def bench_uniform(args) -> int:
    """Uniform-H (shared bases) and H^2 (nested bases) forms of 1/r on
    n = 65536 points on the sphere (leaf 64, eta 2, ACA eps 1e-6, bases cut
    at eps 1e-6), FP64 and mixed-precision storage, beside the H-matrix they
    were converted from.  One step is one y += 2 M x (hmvm_uplan_apply: the
    chain of forward transform, couplings, backward transform and near-field
    passes); the relative error is taken against the oracle's
    uniform::mul_vec / h2::mul_vec on the same matrix, and that restatement
    is timed on one host core beside it."""
    import numpy as np
    import torch

    import oracle
    from hlr_amd import _native as N
    from hlr_amd.hmatrix import INVR_SPHERE, HMatrix, random_vector
    from hlr_amd.plan import Plan
    from hlr_amd.uniform import UPlan, build_uniform

    dev = torch.device("cuda:0")
    n = args.n if args.n != 262144 else 65536
    t0 = time.time()
    hm = HMatrix(INVR_SPHERE.with_(n=n, eps=args.eps))
    forms = {"uniform": build_uniform(hm, args.eps, nested=False), "h2": build_uniform(hm, args.eps, nested=True)}
    log(f"sphere n={n}: H-matrix and its uniform-H / H^2 forms built in {time.time() - t0:.1f} s")
    x_h = random_vector(n, 2)
    x = torch.as_tensor(x_h, device=dev)
    stream = torch.cuda.current_stream()

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    y = torch.zeros_like(x)
    ph = Plan.from_hmatrix(hm, precision="fp64")
    ws = torch.empty(ph.workspace_bytes(N.OP_NORMAL, 1) // 8 + 1, dtype=torch.float64, device=dev)
    ms_h = timed(lambda: ph.apply_ws(2.0, N.OP_NORMAL, x, y, ws), args.steps, args.warmup)
    h_bytes = ph.info()["datasize"]
    ph.close()
    del ws
    res = {}
    for name, um in forms.items():
        yr = np.zeros(n)
        oracle.uniform_mul_vec(um.rowcb, um.colcb, um.blocks, um.arena, 1.0, N.OP_NORMAL, x_h, yr)
        for prec in ("fp64", "valr"):
            up = UPlan(um, precision=prec, eps=args.eps)
            info = up.info()
            ms = timed(lambda: up.apply(2.0, N.OP_NORMAL, x, y), args.steps, args.warmup)
            y1 = torch.zeros_like(x)
            up.apply(1.0, N.OP_NORMAL, x, y1)
            torch.cuda.synchronize()
            err = float(np.linalg.norm(y1.cpu().numpy() - yr) / np.linalg.norm(yr))
            res[f"{name}_{prec}"] = {"ms_per_mvm": ms, "mvm_per_s": 1000.0 / ms, "rel_err_vs_oracle_fp64": err,
                                     "datasize_fp64": info["datasize"],
                                     "hbm_gbs_fp64_bytes": info["datasize"] / (ms * 1e-3) / 1e9,
                                     "levels": info["levels"]}
            log(f"{name} {prec}: {ms:.4f} ms/MVM, {info['datasize'] / 1e6:.1f} MB (FP64), rel.err {err:.2e}")
            up.close()
    cpu = None
    if not args.no_cpu:
        um = forms["h2"]
        yc = np.zeros(n)
        t1 = time.time()
        reps = 0
        while time.time() - t1 < min(args.cpu_seconds, 10.0):
            oracle.uniform_mul_vec(um.rowcb, um.colcb, um.blocks, um.arena, 2.0, N.OP_NORMAL, x_h, yc)
            reps += 1
        tc = time.time() - t1
        cpu = {"value": reps / tc, "unit": "H-MVM/s", "cores": 1, "kind": "port",
               "sample": f"{reps} H^2 MVMs with the oracle's restatement of h2::mul_vec on one thread in {tc:.1f} s"}
    best = res["h2_fp64"]
    out = {
        "metric": "H^2 / uniform-H MVM/s beside the H-matrix they come from (1/r sphere)",
        "value": best["mvm_per_s"], "unit": "H-MVM/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["ms_per_mvm"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded points, generated H-matrix converted to shared / nested bases)",
        "config": {"workload": f"1/r on unit sphere, n={n}, leaf 64, eta 2, ACA eps {args.eps:g}, "
                               f"bases cut at eps {args.eps:g}; H^2 FP64 is the value", "parallelism": "single GPU"},
        "h_matrix_fp64": {"ms_per_mvm": ms_h, "datasize": h_bytes},
        "forms": res,
        "roofline": {"bound": "hbm", "kernel": "whole H^2 product (chain of device passes)",
                     "achieved": best["hbm_gbs_fp64_bytes"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": best["hbm_gbs_fp64_bytes"] / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": best["datasize_fp64"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_batch_baseline(hms, nrhs) -> dict:
    """cpu_baseline leg of the batch: the reference applies one matrix and
    one right-hand side at a time (programs/mvm/mvm.hh:151-204), so every
    member is timed once with the oracle's restatement of tbb::mul_vec_chunk
    on all cores of the affinity set (at most 16), and a step is the sum of
    nrhs times that time over the members (one FP64 H-MVM sampled per member)."""
    import numpy as np

    import oracle
    from oracle.build import CFLAGS
    from hlr_amd.hmatrix import random_vector

    cores = min(len(os.sched_getaffinity(0)), 16)
    step = 0.0
    for i, (hm, r) in enumerate(zip(hms, nrhs)):
        xv = random_vector(hm.n, 9000 + i)
        t = oracle.time_chunk(hm.blocks, hm.arena, xv, np.zeros(hm.n), cores, 1)
        step += r * t
    products = float(sum(nrhs))
    log(f"cpu baseline (batch): {step:.2f} s per step on {cores} threads")
    return {"value": products / step, "unit": "H-MVM/s", "cores": cores, "kind": "port",
            "sample": (f"one FP64 H-MVM per member ({len(hms)}) with the oracle's restatement of tbb::mul_vec_chunk on "
                       f"{cores} threads, scaled by each member's nrhs: {step:.2f} s per step; "
                       f"gcc {' '.join(CFLAGS)}")}


def cpu_baseline(blocks, arena, x_h, n: int, seconds: float) -> dict:
    """cpu_baseline leg: the oracle's restatement of tbb::mul_vec_chunk
    (parallel leaf loop, 64-entry locked chunks of y) timed on the host, on
    all cores of the affinity set (at most 16) and on one core, each on a
    bounded sample of about `seconds` of work.  The oracle is the checker;
    here it is only timed."""
    import numpy as np

    import oracle
    from oracle.build import CFLAGS

    cores = min(len(os.sched_getaffinity(0)), 16)
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    y = np.zeros(n)
    rate, sample = {}, []
    for nt in (cores, 1):
        # calibrate on doubling samples of at least 0.2 s, then time the bounded sample
        nm, t1 = 1, oracle.time_chunk(blocks, arena, x_h, y, nt, 1)
        while t1 < 0.2 and nm < 1 << 20:
            nm *= 2
            t1 = oracle.time_chunk(blocks, arena, x_h, y, nt, nm)
        nmvm = int(seconds * nm / max(t1, 1e-9))
        if nmvm > nm:
            tc = oracle.time_chunk(blocks, arena, x_h, y, nt, nmvm)
        else:
            nmvm, tc = nm, t1
        rate[nt] = nmvm / tc
        sample.append(f"{nmvm} MVMs in {tc:.1f} s on {nt} thread{'s' if nt > 1 else ''}")
        log(f"cpu baseline: {rate[nt]:.3f} MVM/s on {nt} thread(s)")
    return {"value": rate[cores], "unit": "H-MVM/s", "cores": cores, "kind": "port",
            "value_1core": rate[1],
            "sample": (f"FP64 H-MVMs of the same matrix with the oracle's restatement of tbb::mul_vec_chunk "
                       f"(64-entry locked chunks): {'; '.join(sample)}; gcc {' '.join(CFLAGS)}; {model}")}


def bench_grad(args) -> int:
    """Differentiation workload: Matern nu = 1.5, l = 0.1 on n = 131072
    points in the unit cube, TLR tiles of 512, ACA eps 1e-6, FP32 factors;
    L = 1/2 |A x - b|^2 with x, b ~ N(0, 1).  One step is the forward
    product r = A x - b (hmvm_apply_ws) and the adjoint (hmvm_grad):
    gx = A^T r and the FP64 gradient of every stored factor value, with
    V^H x recomputed.  Gradient arenas and workspace are allocated before
    the timed steps."""
    import ctypes as C

    import numpy as np
    import torch

    from hlr_amd import _native as N
    from hlr_amd.hmatrix import MATERN_TLR, HMatrix, random_vector
    from hlr_amd.plan import Plan, _stream_handle

    dev = torch.device("cuda:0")
    t0 = time.time()
    hm = HMatrix(MATERN_TLR)
    log(f"matrix: n={hm.n}, {len(hm.blocks)} tiles, built in {time.time() - t0:.1f} s")
    plan = Plan.from_hmatrix(hm, precision="fp32")
    info = plan.info()
    x = torch.as_tensor(random_vector(hm.n, 41, gaussian=True), device=dev)
    b = torch.as_tensor(random_vector(hm.n, 42, gaussian=True), device=dev)
    r = torch.empty_like(x)
    gx = torch.zeros_like(x)
    elems = [plan.arena_elems(f) for f in range(N.NFMT)]
    garena = [torch.zeros(e, dtype=torch.float64, device=dev) if e else None for e in elems]
    ptrs = (C.c_void_p * N.NFMT)(*[C.c_void_p(g.data_ptr()) if g is not None else None for g in garena])
    ws = torch.empty(plan.workspace_bytes(N.OP_NORMAL, 1) // 8 + 1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sh = _stream_handle(stream)

    def forward():
        r.copy_(b).neg_()
        plan.apply_ws(1.0, N.OP_NORMAL, x, r, ws)          # r = A x - b

    def adjoint():
        gx.zero_()
        N.check(N.hmvm().hmvm_grad(plan.handle, C.c_void_p(x.data_ptr()), C.c_void_p(r.data_ptr()),
                                   C.c_void_p(gx.data_ptr()), ptrs, N.F64, sh), "hmvm_grad")

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    steps, warmup = min(args.steps, 20), min(args.warmup, 3)
    ms_step = timed(lambda: (forward(), adjoint()), steps, warmup)
    ms_fwd = timed(forward, steps, warmup)
    forward()
    ms_adj = timed(adjoint, steps, warmup)
    loss = 0.5 * float(torch.dot(r, r))
    # algorithmic bytes of the adjoint: the V factors (V^H x recomputed), all
    # factors again (A^T r), 8 B written per stored value, plus the vectors
    n = hm.n
    adj_bytes = info["bytes_vpass"] + info["factor_bytes"] + 8.0 * sum(elems) + 8.0 * 4 * n
    out = {
        "metric": "loss + gradient steps/s (L = 1/2 |Ax - b|^2, gradients w.r.t. x and every factor)",
        "value": 1000.0 / ms_step, "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factors, f64 vectors and gradients",
        "data": "synthetic (seeded points, generated TLR matrix, x and b ~ N(0,1))",
        "config": {"workload": f"Matern nu=1.5 l=0.1, n={n} in the unit cube, TLR tile 512, ACA eps 1e-6, FP32",
                   "parallelism": "single GPU"},
        "forward_ms": ms_fwd, "adjoint_ms": ms_adj, "loss": loss,
        "factor_bytes": info["factor_bytes"], "gradient_values": int(sum(elems)),
        "adjoint_algorithmic_bytes": adj_bytes, "adjoint_gbs": adj_bytes / (ms_adj * 1e-3) / 1e9,
        "forward_gbs": info["datasize"] / (ms_fwd * 1e-3) / 1e9,
        "cpu_baseline": None,
    }
    log(f"step {ms_step:.3f} ms: forward {ms_fwd:.3f} ms, adjoint {ms_adj:.3f} ms "
        f"({out['adjoint_gbs']:.0f} GB/s algorithmic)")
    print(json.dumps(out), flush=True)
    return 0


def bench_batch(args) -> int:
    """64 Matern (nu = 1.5) H-matrices, l ~ U[0.05, 0.3], n log-uniform in
    [2048, 32768], nrhs uniform in {1, 4, 16, 64}, FP64 factors, X ~ N(0,1).
    One step = the forward products of all 64 members (a fixed sequence of
    stream-ordered kernels, schedule picked by the autotuner); the same products issued
    plan by plan and the graph replay are timed beside it."""
    import numpy as np
    import torch

    from hlr_amd import _native as N
    from hlr_amd.batch import Batch
    from hlr_amd.hmatrix import HMatrix, batch_problems, random_vector
    from hlr_amd.plan import Plan

    dev = torch.device("cuda:0")
    t0 = time.time()
    probs = batch_problems(count=64, seed=7)
    hms = [HMatrix(p) for p, _ in probs]
    nrhs = [r for _, r in probs]
    log(f"batch: 64 matrices, n {min(h.n for h in hms)}..{max(h.n for h in hms)}, "
        f"nrhs counts {dict(zip(*np.unique(nrhs, return_counts=True)))}, built in {time.time() - t0:.1f} s")
    plans = [Plan.from_hmatrix(h, precision="fp64") for h in hms]
    infos = [p.info() for p in plans]
    xs, ys = [], []
    for i, (h, r) in enumerate(zip(hms, nrhs)):
        X = np.stack([random_vector(h.n, 1000 * i + j, gaussian=True) for j in range(r)])
        xs.append(torch.as_tensor(X if r > 1 else X[0], device=dev))
        ys.append(torch.zeros_like(xs[-1]))
    b = Batch(plans, nrhs)
    chosen = b.tune(xs, ys)
    stream = torch.cuda.current_stream()

    def timed(fn, steps, warmup, st=stream):
        # events on the stream the work is issued on
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(steps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    ms_batch = timed(lambda: b.apply(xs, ys), args.steps, args.warmup)
    wss = [torch.empty(p.workspace_bytes(N.OP_NORMAL, r) // 8 + 1, dtype=torch.float64, device=dev)
           for p, r in zip(plans, nrhs)]

    def per_plan():
        for p, x, y, w in zip(plans, xs, ys, wss):
            p.apply_ws(1.0, N.OP_NORMAL, x, y, w)

    ms_loop = timed(per_plan, args.steps, args.warmup)
    # graph capture needs a stream of its own (not the legacy default stream)
    gs = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(gs):
        b.capture(xs, ys, stream=gs)
    gs.synchronize()
    ms_graph = timed(lambda: b.replay(stream=gs), args.steps, args.warmup, st=gs)
    b.check(stream=gs)
    products = float(sum(nrhs))
    cpu = None if args.no_cpu else cpu_batch_baseline(hms, nrhs)
    # factor bytes once per member and call, vectors per right-hand side
    nbytes = float(sum(i["datasize"] + 8.0 * 2 * (r - 1) * (i["nrows"] + i["ncols"]) for i, r in zip(infos, nrhs)))
    flops = float(sum(i["flops"] * r for i, r in zip(infos, nrhs)))
    log(f"batch: {ms_batch:.3f} ms/step (graph replay {ms_graph:.3f}, plan by plan {ms_loop:.3f}); "
        f"schedule {chosen}; {nbytes / 1e9:.3f} GB and {flops / 1e9:.2f} GFlop per step")
    # HBM traffic of the batched kernels from the committed rocprofv3 PMC
    # summary of this workload (FETCH_SIZE + WRITE_SIZE per step), when it exists
    traffic = None
    tfile = ROOT / "profiles" / "traffic_batch_matern64.txt"
    if tfile.exists():
        traffic = 0.0
        for line in tfile.read_text().splitlines():
            f = line.split()
            if len(f) >= 4 and not line.startswith("#") and f[0] == "batch":
                traffic += float(f[-2]) + float(f[-1])
    out = {
        "metric": "batched H-MVM/s (64 Matern H-matrices, mixed nrhs)",
        "value": products / (ms_batch * 1e-3),
        "unit": "H-MVM/s (one right-hand side = one product)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_batch,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded points, generated H-matrices, X ~ N(0,1))",
        "config": {"workload": "64 Matern nu=1.5 H-matrices, l~U[0.05,0.3], n log-uniform [2048,32768], "
                               "nrhs in {1,4,16,64}, leaf 64, eta 2, ACA eps 1e-6, FP64",
                   "parallelism": "single GPU"},
        "ms_per_step_graph_replay": ms_graph,
        "ms_per_step_plan_by_plan": ms_loop,
        "schedule": chosen,
        "hbm_gbs": nbytes / (ms_batch * 1e-3) / 1e9,
        "gflops": flops / (ms_batch * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": "whole step (batched kernels)",
                     "achieved": nbytes / (ms_batch * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": nbytes / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": (f"profiles/{tfile.name} (rocprofv3 PMC FETCH_SIZE + WRITE_SIZE of the batched "
                                        "kernels on this workload, copy-calibrated; not measured in this run)")
                     if traffic is not None else None,
                     "algorithmic_bytes": nbytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    b.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
