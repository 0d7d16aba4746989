#!/usr/bin/env python3
"""HODLR factor+solve benchmark on one MI355X.

Workload (the reference's hot path at its GPU-sized configuration):
N = 2^18 = 262144, binary tree of 512 leaves of size 512 (9 levels), every
off-diagonal block at every level U V^T with rank 64, U, V iid N(0,1)/8,
leaf blocks G + 512 I (G iid N(0,1)), 8 right-hand sides iid U[0,1), seed 0,
FP64.  One step = one fused factor + solve of the whole system
(hodlr_factor_solve), inputs resident in HBM before timing starts; K is only
read (the leaf factors go to the plan), each step solves its own pre-staged
copy of B (overwritten by the solution).

Prints ONE JSON line on stdout; diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0       # MI355X HBM3E
LD_PAD = 8                  # leading dimension N + 8 for every N-row operand


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_alg(N, m, r, s, L=None):
    """Algorithmic FLOPs of one factor+solve with the reference's column
    layout (leading terms; L = log2(N/m) levels unless given, t = 0 is the
    root; N/m leaves)."""
    P = N // m
    if L is None:
        L = int(round(math.log2(P)))
    f = P * (2.0 / 3.0) * m ** 3 + 2.0 * P * m * m * r
    for t in range(L):
        f += 2 * N * r * r + 4 * N * r * r * t + 2 ** t * (16.0 / 3.0) * r ** 3 + 2 ** t * 8 * r ** 3 * t
    f += 2.0 * N * m * s
    for t in range(L):
        f += 4.0 * N * r * s + 2 ** t * 8.0 * r * r * s
    return f


def flops_exec(N, m, r, s):
    """FLOPs this library executes (node-sum level loop): leaf LU, leaf solve
    on [U | B], leaf projections, per node a 2r x 2r core LU and solve on
    r + s columns, the top-down pass (L products r x r by r x s per leaf)
    and the final B -= Y G."""
    P = N // m
    L = int(round(math.log2(P)))
    nc = r + s
    f = P * (2.0 / 3.0) * m ** 3 + 2.0 * P * m * m * (nc if L > 0 else s)
    if L > 0:
        f += 2.0 * N * r * nc                                            # projections
        f += (P - 1) * ((2.0 / 3.0) * (2 * r) ** 3 + 2.0 * (2 * r) ** 2 * nc)  # cores
        f += P * L * 2.0 * r * r * s                                     # top-down pass
        f += 2.0 * N * r * s                                             # B -= Y G
    return f


KERNELS_SRC = os.path.join(ROOT, "fastsolver2_amd", "csrc", "hodlr_kernels.hip")
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def kernels_source_sha256():
    with open(KERNELS_SRC, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def pmc_mean_kib(d, counter, kernel_substr):
    """Mean per-dispatch value of a rocprofv3 --pmc counter over the
    dispatches of kernels whose name contains kernel_substr, from the
    counter_collection CSVs under d (values summed per dispatch)."""
    per = {}
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter or kernel_substr not in row.get("Kernel_Name", ""):
                    continue
                key = (path, row.get("Dispatch_Id"))
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    if not per:
        raise SystemExit(f"no {counter} rows for '{kernel_substr}' under {d}")
    return sum(per.values()) / len(per), len(per)


def write_traffic(dirs, kernel_substr, workload, copy_bytes):
    """profiles/pmc_traffic.json from four rocprofv3 --pmc runs: FETCH_SIZE
    and WRITE_SIZE of the kernel, and of a plain device copy of copy_bytes
    (the calibration: the copy reads and writes copy_bytes exactly).  The
    JSON carries the SHA-256 of the kernel source it was measured on; the
    bench reports the traffic only while the source is unchanged."""
    fdir, wdir, cfdir, cwdir = dirs
    fk, nf = pmc_mean_kib(fdir, "FETCH_SIZE", kernel_substr)
    wk, nw = pmc_mean_kib(wdir, "WRITE_SIZE", kernel_substr)
    cf, _ = pmc_mean_kib(cfdir, "FETCH_SIZE", "copyBuffer")
    cw, _ = pmc_mean_kib(cwdir, "WRITE_SIZE", "copyBuffer")
    fscale = copy_bytes / 1024.0 / cf
    wscale = copy_bytes / 1024.0 / cw
    rd, wr = fk * 1024.0 * fscale, wk * 1024.0 * wscale
    N, m = workload["n"], workload["leaf"]
    out = {"kernel": "getrf_leaf", "kernel_match": kernel_substr,
           "workload": f"N={N},leaf={m},rank={workload['rank']},nrhs={workload['nrhs']}",
           "source_sha256": kernels_source_sha256(), "dispatches": [nf, nw],
           "fetch_kib_raw": fk, "write_kib_raw": wk, "copy_probe_bytes": copy_bytes,
           "copy_fetch_kib": cf, "copy_write_kib": cw, "fetch_scale": fscale,
           "write_scale": wscale, "hbm_read_bytes_per_launch": rd,
           "hbm_write_bytes_per_launch": wr, "hbm_bytes_per_launch": rd + wr,
           "algorithmic_bytes_per_launch": 16 * N * m}
    with open(TRAFFIC_JSON, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def host_threads():
    """CPUs this process may run on: its affinity set, capped by
    OMP_NUM_THREADS when that is set (the host's CPU share per GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(K, U, V, B, m, nthreads, runs=5):
    """The reference algorithm (oracle restatement of hsolve + the level loop)
    with the best CPU BLAS found, on the same arrays."""
    import numpy as np
    from oracle import oracle as O
    name = O.use_openblas()
    N = K.shape[0]
    L = int(round(math.log2(N // m)))
    tl = min(L, int(math.floor(math.log2(max(1, nthreads)))))
    times = []
    for i in range(runs + 1):
        t0 = time.perf_counter()
        O.hodlr_solve(K, U, V, B, m, task_level=tl, nthreads=nthreads)
        dt = time.perf_counter() - t0
        if i > 0:
            times.append(dt)
        log(f"  cpu run {i}: {dt:.3f} s ({'warm-up' if i == 0 else 'timed'})")
    return statistics.median(times), name, tl


def bench_config4(F, torch, W, log):
    """N = 2^20, 2048 leaves of 512 (11 levels), rank 64, 4 RHS iid N(0,1),
    leaf blocks G + s I with s calibrated to cond ~ 1e4: mixed + refinement
    to 1e-12 next to the plain FP64 factor + solve, one timed run each."""
    from fastsolver2_amd.workloads import conditioned_shift_global
    N, m, r, s = 1 << 20, 512, 64, 4
    L = 11
    shift, cnd = conditioned_shift_global(N, m, r, 1e4, seed=0)
    K, U, V, B = F.workload_gaussian(N, m, r, s, seed=0, shift=shift, rhs="normal", pad=LD_PAD)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    mplan = F.HodlrPlan(N, m, L, r, nrhs_max=s, precision=F.MIXED_IR)
    X = F.fortran_empty(N, s, pad=LD_PAD)
    tm = []
    for i in range(W + 1):
        torch.cuda.synchronize()
        ev0.record()
        X.copy_(B)
        mplan.factor_solve(K, U, V, X)
        X, st = mplan.solve_ir(K, U, V, B, X=X, tol=1e-12, maxit=30, x_given=True)
        ev1.record()
        torch.cuda.synchronize()
        tm.append(ev0.elapsed_time(ev1))
    mplan.close()
    plan = F.HodlrPlan(N, m, L, r, nrhs_max=s)
    Kf = F.fortran_clone(K)
    Xf = F.fortran_clone(B)
    tf = []
    for i in range(W + 1):
        Kf.copy_(K)
        Xf.copy_(B)
        torch.cuda.synchronize()
        ev0.record()
        plan.factor_solve(Kf, U, V, Xf)
        ev1.record()
        torch.cuda.synchronize()
        tf.append(ev0.elapsed_time(ev1))
    plan.close()
    err = (X - Xf).abs().max().item() / Xf.abs().max().item()
    out = {"workload": f"N={N} leaf={m} levels={L} rank={r} nrhs={s}, leaf shift {shift:.4f} "
                       f"(estimated cond_2(A) {cnd:.3g}), B iid N(0,1)",
           "mixed_ir_ms": round(tm[-1], 3), "fp64_ms": round(tf[-1], 3),
           "speedup_vs_fp64": round(tf[-1] / tm[-1], 3), "iterations": st["iterations"],
           "residual": st["residual"], "max_rel_diff_vs_fp64": err}
    log(f"config N=2^20 cond~1e4: mixed+IR {tm[-1]:.2f} ms ({st['iterations']} corrections, "
        f"residual {st['residual']:.2e}), FP64 {tf[-1]:.2f} ms")
    del K, U, V, B, X, Kf, Xf
    return out


def bench_batch(F, torch, W, K_, log, nsys=256):
    """256 independent systems (N log-uniform in [4096, 65536] rounded to the
    leaf, leaf in {128, 256, 512}, rank uniform in [16, 100], 1 RHS U[0,1),
    seed 0), P leaves each in a tree of 2^ceil(log2 P) leaf slots (the
    trailing slots empty); autotuned, one graph per batch."""
    from fastsolver2_amd.workloads import batch_operands, batch_systems
    systems = batch_systems(nsys, seed=0)
    Ks, Us, Vs, Bs = batch_operands(systems, seed=0, pad=LD_PAD)
    B0 = [F.fortran_clone(x) for x in Bs]
    b = F.HodlrBatch(systems)
    t0 = time.perf_counter()
    lanes, tile, tune_ms = b.tune(Ks, Us, Vs, Bs)
    tune_s = time.perf_counter() - t0
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ts = []
    for i in range(W + K_):
        for x, x0 in zip(Bs, B0):  # B is overwritten by the solution; K is only read
            x.copy_(x0)
        torch.cuda.synchronize()
        ev0.record()
        b.factor_solve(Ks, Us, Vs, Bs)
        ev1.record()
        torch.cuda.synchronize()
        if i >= W:
            ts.append(ev0.elapsed_time(ev1))
    st = b.status()
    if st[0] != 0:
        log(f"batch status {st}")
    worst = 0.0
    for sy, K, U, V, X, B in zip(systems, Ks, Us, Vs, Bs, B0):
        R = F.matvec(K, U, V, X, sy["leaf"], alpha=-1.0)
        R.add_(B)
        worst = max(worst, (R.abs().max() / B.abs().max()).item())
    ms = statistics.median(ts)
    fl = sum(flops_alg(sy["n"], sy["leaf"], sy["rank"], 1, sy["levels"]) for sy in systems)
    out = {"workload": f"{nsys} systems, N log-uniform [4096, 65536] (sum {sum(s['n'] for s in systems)}"
                       f" rows, no padding: empty trailing leaf slots), leaf {{128,256,512}}, "
                       "rank U[16,100], 1 RHS, seed 0, one HIP graph",
           "ms": round(ms, 3), "systems_per_s": round(nsys * 1000.0 / ms, 1),
           "batches_per_s": round(1000.0 / ms, 3), "gflops_alg": round(fl / ms / 1e6, 1),
           "tuned": {"lanes": lanes, "tile": tile, "tune_replay_ms": round(tune_ms, 3),
                     "tune_seconds": round(tune_s, 2)},
           "max_relative_residual": worst}
    log(f"batch of {nsys}: {ms:.2f} ms ({nsys * 1000.0 / ms:.0f} systems/s, {fl / ms / 1e6:.0f} GFLOP/s), "
        f"tuned lanes={lanes} tile={tile}, worst residual {worst:.2e}")
    b.close()
    del Ks, Us, Vs, Bs, B0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device 4 times (calibrates HBM counters)")
    ap.add_argument("--config4", action=argparse.BooleanOptionalAction, default=True,
                    help="mixed precision at N=2^20, cond~1e4 (default on)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--write-traffic", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPYF_DIR",
                                                         "COPYW_DIR"),
                    help="write profiles/pmc_traffic.json from rocprofv3 --pmc output dirs "
                         "(no GPU work)")
    ap.add_argument("--traffic-kernel", default="k_getrf_ll<4, double>")
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--leaf", type=int, default=512)
    ap.add_argument("--rank", type=int, default=64)
    ap.add_argument("--nrhs", type=int, default=8)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the HODLR path is "
              "built and measured on a single MI355X, multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)

    if args.write_traffic:
        write_traffic(args.write_traffic, args.traffic_kernel,
                      {"n": args.n, "leaf": args.leaf, "rank": args.rank, "nrhs": args.nrhs},
                      args.copy_probe or (1 << 30))
        return

    import numpy as np
    import torch
    import fastsolver2_amd as F

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.copy_probe:
        n = args.copy_probe // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 4 device copies of {n * 8} bytes")
        return
    # all work on one non-default stream (default priority): the library
    # captures its refinement solve as a HIP graph on such a stream
    side = torch.cuda.Stream()
    torch.cuda.set_stream(side)
    N, m, r, s = args.n, args.leaf, args.rank, args.nrhs
    L = int(round(math.log2(N // m)))
    assert m << L == N
    W, K_ = args.warmup, args.steps

    log(f"workload: N={N} leaf={m} levels={L} rank={r} nrhs={s} FP64; steps={K_} warmup={W}")
    # operands with leading dimension N + 8: a power-of-two column stride
    # makes a leaf's columns alias in the HBM channel interleave
    K0, U, V, B0 = F.workload_gaussian(N, m, r, s, seed=0, pad=LD_PAD)
    NB_PROF = 2  # steps of the per-kernel breakdown pass
    Bs = [F.fortran_clone(B0) for _ in range(W + K_ + NB_PROF)]
    plan = F.HodlrPlan(N, m, L, r, nrhs_max=s)
    log(f"plan device memory: {plan.device_bytes / 2**30:.2f} GiB")
    torch.cuda.synchronize()

    for i in range(W):
        plan.factor_solve(K0, U, V, Bs[i])
    torch.cuda.synchronize()

    # timed region: events only around the leaf LU launches (the roofline
    # kernel); every recorded launch adds two event records to the stream
    F.profile_reset()
    F.profile_enable(True, kinds=["getrf_leaf"])
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for i in range(K_):
        plan.factor_solve(K0, U, V, Bs[W + i])
    ev1.record()
    torch.cuda.synchronize()
    F.profile_enable(False)
    ms = ev0.elapsed_time(ev1) / K_
    prof = F.profile_read()
    st, where = plan.status()
    if st != 0:
        log(f"factorization status {st} at {where}")
        sys.exit(1)

    # correctness guard: residual of the last timed solve with the device matvec
    X = Bs[W + K_ - 1]
    R = F.matvec(K0, U, V, X, m, alpha=-1.0)
    R.add_(B0)
    torch.cuda.synchronize()
    rel_res = (R.abs().max() / B0.abs().max()).item()
    log(f"relative residual max|b - A x| / max|b| = {rel_res:.3e}")

    # factor-reuse solve rate (same factorization, fresh right-hand sides)
    Bsol = [F.fortran_clone(B0) for _ in range(K_)]
    torch.cuda.synchronize()
    ev0.record()
    for i in range(K_):
        plan.solve(Bsol[i])
    ev1.record()
    torch.cuda.synchronize()
    solve_ms = ev0.elapsed_time(ev1) / K_
    del Bsol

    # per-kernel breakdown: a separate pass with every launch recorded
    F.profile_reset()
    F.profile_enable(True)
    for i in range(NB_PROF):
        plan.factor_solve(K0, U, V, Bs[W + K_ + i])
    torch.cuda.synchronize()
    F.profile_enable(False)
    prof_all = F.profile_read()

    fa = flops_alg(N, m, r, s)
    fe = flops_exec(N, m, r, s)
    value = 1000.0 / ms
    log(f"factor+solve: {ms:.3f} ms/solve = {value:.1f} solves/s; "
        f"{fa / ms / 1e6:.1f} GFLOP/s (reference-layout count {fa / 1e9:.1f} GF), "
        f"{fe / ms / 1e6:.1f} GFLOP/s executed ({fe / 1e9:.1f} GF); solve-only {solve_ms:.3f} ms")

    # per-kernel table (breakdown pass) and roofline of the dominant kernel
    # class (its launches timed inside the timed region)
    log(f"per-kernel breakdown ({NB_PROF} profiled steps, every launch bracketed by events):")
    log(f"{'kernel':12s} {'launches':>8s} {'ms/step':>9s} {'avg us':>9s} {'TFLOP/s':>8s} {'GB/s':>8s}")
    dom, dom_ms = None, -1.0
    for k, d in prof_all.items():
        if d["count"] == 0:
            continue
        per_step = d["ms"] / NB_PROF
        avg_us = 1000.0 * d["ms"] / d["count"]
        tf = d["flops"] / (d["ms"] * 1e9) if d["ms"] > 0 else 0.0
        gbs = d["bytes"] / (d["ms"] * 1e6) if d["ms"] > 0 else 0.0
        log(f"{k:12s} {d['count']:8d} {per_step:9.3f} {avg_us:9.1f} {tf:8.2f} {gbs:8.0f}")
        if d["ms"] > dom_ms:
            dom, dom_ms = k, d["ms"]
    d = prof[dom] if prof[dom]["count"] > 0 else prof_all[dom]
    avg_s = d["ms"] / d["count"] / 1000.0
    fl_per = d["flops"] / d["count"]
    by_per = d["bytes"] / d["count"]
    intensity = fl_per / by_per if by_per > 0 else float("inf")
    if intensity >= PEAK_FP64_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9):
        ach = fl_per / avg_s / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / PEAK_FP64_TFLOPS, 4)}
    else:
        ach = by_per / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4)}
    roof["kernel"] = dom
    roof["avg_launch_us"] = round(avg_s * 1e6, 2)
    roof["algorithmic_flops_per_launch"] = fl_per
    roof["algorithmic_bytes_per_launch"] = by_per
    # HBM traffic per launch from the calibrated --pmc passes, valid only
    # for the kernel source it was measured on (null otherwise)
    traffic = None
    if os.path.exists(TRAFFIC_JSON):
        try:
            tj = json.load(open(TRAFFIC_JSON))
            if (tj.get("kernel") == dom
                    and tj.get("workload") == f"N={N},leaf={m},rank={r},nrhs={s}"
                    and tj.get("source_sha256") == kernels_source_sha256()):
                traffic = tj.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    roof["traffic"] = traffic

    # mixed precision + FP64 refinement to 1e-12 on the same system: FP32
    # leaf LU, FP64 levels, residuals with the FP64 HODLR matvec
    mixed = None
    if not args.no_mixed:
        mplan = F.HodlrPlan(N, m, L, r, nrhs_max=s, precision=F.MIXED_IR)
        Xm = F.fortran_empty(N, s, pad=LD_PAD)
        mts, sts = [], None
        for i in range(W + K_):
            torch.cuda.synchronize()
            ev0.record()
            Xm.copy_(B0)
            mplan.factor_solve(K0, U, V, Xm)  # K0 untouched: the FP32 copy is factored
            Xm, sts = mplan.solve_ir(K0, U, V, B0, X=Xm, tol=1e-12, maxit=10, x_given=True)
            ev1.record()
            torch.cuda.synchronize()
            if i >= W:
                mts.append(ev0.elapsed_time(ev1))
        mms = statistics.median(mts)
        err = (Xm - Bs[W + K_ - 1]).abs().max().item() / Bs[W + K_ - 1].abs().max().item()
        mixed = {"workload": f"same system as value (N={N} leaf={m} rank={r} nrhs={s}), "
                             "FP32 leaf LU + FP64 levels, FP64 refinement to 1e-12",
                 "value": round(1000.0 / mms, 3), "unit": "solves/s", "ms": round(mms, 4),
                 "iterations": sts["iterations"], "residual": sts["residual"],
                 "history": sts["history"], "max_rel_diff_vs_fp64": err,
                 "speedup_vs_fp64": round(ms / mms, 3),
                 "gflops_alg": round(fa / mms / 1e6, 2)}
        log(f"mixed + IR: {mms:.3f} ms/solve ({1000.0 / mms:.1f} solves/s, {ms / mms:.2f}x FP64), "
            f"{sts['iterations']} corrections, residual {sts['residual']:.2e}, "
            f"max rel diff vs FP64 {err:.2e}")
        mplan.close()
        del Xm
        if args.config4:
            mixed["config4"] = bench_config4(F, torch, W, log)

    # differentiable solve (N=65536, leaf 256, rank 32, 16 RHS): forward
    # factor+solve, then the adjoint solve and the K/U/V gradients
    adj = None
    if not args.no_adjoint:
        aN, am, ar, as_ = 65536, 256, 32, 16
        aL = int(round(math.log2(aN // am)))
        aK0, aU, aV, aB0 = F.workload_gaussian(aN, am, ar, as_, seed=0, pad=LD_PAD)
        aplan = F.HodlrPlan(aN, am, aL, ar, nrhs_max=as_)
        aK = F.fortran_clone(aK0)
        aX = F.fortran_clone(aB0)
        Xstar = F.fortran(torch.randn(as_, aN, dtype=torch.float64, device="cuda").t())
        fw = []
        for i in range(W + K_):
            aK.copy_(aK0)
            aX.copy_(aB0)
            ev0.record()
            aplan.factor_solve(aK, aU, aV, aX)
            ev1.record()
            torch.cuda.synchronize()
            if i >= W:
                fw.append(ev0.elapsed_time(ev1))
        G = aX - Xstar
        Lam = F.fortran_empty(aN, as_)
        for i in range(W):
            Lam.copy_(G)
            aplan.solve_adjoint(Lam)
            F.grad(aU, aV, aX, Lam, am)
        torch.cuda.synchronize()
        bw = []
        for i in range(K_):
            Lam.copy_(G)
            torch.cuda.synchronize()
            ev0.record()
            aplan.solve_adjoint(Lam)
            gK, gU, gV = F.grad(aU, aV, aX, Lam, am)
            ev1.record()
            torch.cuda.synchronize()
            bw.append(ev0.elapsed_time(ev1))
        assert aplan.status()[0] == 0
        adj = {"workload": f"N={aN} leaf={am} levels={aL} rank={ar} nrhs={as_} fp64, "
                           "loss 0.5||x - x*||^2",
               "adjoint_ms": round(statistics.median(bw), 4),
               "forward_ms": round(statistics.median(fw), 4)}
        log(f"adjoint (transposed solve + dK, dU, dV): {adj['adjoint_ms']:.3f} ms; "
            f"forward factor+solve {adj['forward_ms']:.3f} ms")
        aplan.close()
        del aK, aK0, aU, aV, aB0, aX, Lam, gK, gU, gV

    batch = None
    if not args.no_batch:
        batch = bench_batch(F, torch, W, K_, log)

    cpu = None
    if not args.no_cpu:
        nthreads = args.cpu_threads or host_threads()
        log(f"cpu baseline: reference algorithm (oracle), {nthreads} threads, full workload")
        Kh, Uh, Vh, Bh = (np.asfortranarray(t.cpu().numpy()) for t in (K0, U, V, B0))
        t_cpu, blas, tl = cpu_baseline(Kh, Uh, Vh, Bh, m, nthreads)
        cpu = {"value": round(1.0 / t_cpu, 4), "unit": "solves/s", "cores": nthreads,
               "kind": "port",
               "sample": f"full workload N={N} leaf={m} rank={r} nrhs={s}; restated hsolve + "
                         f"level loop, {2 ** tl} leaf tasks on {nthreads} threads (every CPU "
                         f"of this process's affinity set, capped by OMP_NUM_THREADS; "
                         f"{os.cpu_count()} CPUs visible), BLAS {blas}; median of 5 after 1 "
                         "warm-up",
               "seconds_per_solve": round(t_cpu, 4),
               "gflops_alg": round(fa / t_cpu / 1e9, 2)}
        log(f"cpu: {t_cpu:.3f} s/solve -> GPU speed-up {t_cpu * 1000.0 / ms:.0f}x")

    out = {
        "metric": "HODLR factor+solve throughput, N=2^18 leaf 512 rank 64 8 RHS, FP64",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K_,
        "warmup": W,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device Philox4x64-10, seed 0)",
        "config": {"workload": f"hodlr_factor_solve N={N} leaf={m} levels={L} rank={r} "
                               f"nrhs={s} fp64", "n": N, "leaf": m, "levels": L, "rank": r,
                   "nrhs": s},
        "gflops_alg": round(fa / ms / 1e6, 2),
        "gflops_exec": round(fe / ms / 1e6, 2),
        "solve_only_ms": round(solve_ms, 4),
        "relative_residual": rel_res,
        "mixed_ir": mixed,
        "adjoint_ms": adj["adjoint_ms"] if adj else None,
        "adjoint": adj,
        "batch": batch,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
