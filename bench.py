#!/usr/bin/env python3
"""bench.py -- DGETRF + DGETRS throughput on one MI355X.

Workload (one "step"): the plain FP64 LU solve at n=16384, nrhs=16, A and B
i.i.d. N(0,1) from the library's seeded generator: DGETRF on A, then
DGETRS('N') on B, both on device-resident inputs through the device entry
points.  Each step works on its own copy of A and B, staged in HBM before
the timed region.

Printed: ONE JSON line with the metric, the GEMM roofline (trailing-update
kernel, FP64 matrix peak) measured with HIP events inside the timed region,
and a CPU baseline: the reference DGETRF + reference BLAS built from the
reference sources (oracle/_ref) -- or its C++ restatement where that build
is absent -- timed on one host core on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time-to-FP64-backward-error solve, MP-IR vs FP64 LU, n=32768; DGETRF TFLOP/s")
PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 matrix
PEAK_HBM = 8000.0  # GB/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--nrhs", type=int, default=16)
    p.add_argument("--seed", type=int, default=16384)
    p.add_argument("--cpu-n", type=int, default=2048, help="CPU baseline sample size")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--mp-n", type=int, default=32768, help="order of the mixed-precision comparison")
    p.add_argument("--no-mp", action="store_true", help="skip the mixed-precision comparison")
    p.add_argument("--no-phase", action="store_true",
                   help="time the steps without per-phase events (no roofline / phase fields)")
    p.add_argument("--mp-half", action="store_true",
                   help="add the FP16 / BF16 GMRES-IR variants to the mixed-precision comparison")
    p.add_argument("--extras", action="store_true",
                   help="also time DGEEQU, DGEEQUB, DLAQGE, DGERFS and the adjoint at the benchmark's order")
    p.add_argument("--calib-copy-mib", type=int, default=0,
                   help="run one plain device copy of this many MiB first (calibrates HBM byte counters)")
    return p.parse_args()


def cpu_baseline(n_sample, seed):
    """Reference DGETRF (flang build of the reference sources, 1 core) on an
    n_sample x n_sample N(0,1) matrix; TFLOP/s of 2/3 n^3."""
    import numpy as np
    import lapack_amd as la
    import oracle
    a = la.matrix(3, seed, n_sample, n_sample)
    kind = "reference" if oracle.ref_available() else "port"
    fn = oracle.ref_dgetrf if kind == "reference" else oracle.dgetrf
    best = None
    for _ in range(2):
        t = time.perf_counter()
        _, _, info = fn(a)
        dt = time.perf_counter() - t
        assert info == 0
        best = dt if best is None else min(best, dt)
    flops = 2.0 / 3.0 * n_sample ** 3
    return {
        "value": flops / best / 1e12,
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": kind,
        "sample": f"DGETRF n={n_sample} N(0,1), reference BLAS, best of 2, {best:.2f} s",
    }


def update_traffic(n):
    """HBM bytes per trailing-update launch measured by the two rocprofv3 PMC
    passes (FETCH_SIZE, WRITE_SIZE, copy-calibrated) summarised in profiles/,
    for the configuration they were taken on; None otherwise."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"dgetrf_n{n}_update_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def extras(n, nrhs, seed):
    """Times (ms, device-resident operands, synchronous entry points, mean of
    three calls after one untimed call) of the routines around the solve at
    order n: DGEEQU and DGEEQUB on A, DLAQGE with their factors, DGERFS('N')
    refining the DGETRS solution of nrhs right-hand sides, and the adjoint
    (Bbar = A^-T Xbar, Abar = -Bbar X^T, an n x n output) on the saved LU."""
    import numpy as np
    import torch
    import lapack_amd as la
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    a = torch.randn(n, n, generator=g, device=dev, dtype=torch.float64)  # column-major: A(i, j) = a[j, i]
    b = torch.randn(nrhs, n, generator=g, device=dev, dtype=torch.float64)
    r = torch.empty(n, device=dev, dtype=torch.float64)
    c = torch.empty(n, device=dev, dtype=torch.float64)

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return round((time.perf_counter() - t0) / reps * 1e3, 3)

    out = {"n": n, "nrhs": nrhs}
    out["dgeequ_ms"] = timed(lambda: la.dgeequ(n, n, a, n, r, c))
    out["dgeequb_ms"] = timed(lambda: la.dgeequ(n, n, a, n, r, c, True))
    info, rowcnd, colcnd, amax = la.dgeequ(n, n, a, n, r, c)
    assert info == 0
    a2 = a.clone()
    out["dlaqge_ms"] = timed(lambda: la.dlaqge(n, n, a2, n, r, c, rowcnd, colcnd, amax))
    del a2
    lu = a.clone()
    ipiv = torch.empty(n, device=dev, dtype=torch.int32)
    assert la.dgetrf(n, n, lu, n, ipiv) == 0
    x = b.clone()
    assert la.dgetrs("N", n, nrhs, lu, n, ipiv, x, n) == 0
    x0 = x.clone()
    ferr = np.zeros(nrhs)
    berr = np.zeros(nrhs)

    def rfs():
        x.copy_(x0)
        assert la.dgerfs("N", n, nrhs, a, n, lu, n, ipiv, b, n, x, n, ferr, berr) == 0
    out["dgerfs_ms"] = timed(rfs)
    out["dgerfs_berr_max"] = float(berr.max())
    xbar = torch.randn(nrhs, n, generator=g, device=dev, dtype=torch.float64)
    bbar = torch.empty(nrhs, n, device=dev, dtype=torch.float64)
    abar = torch.empty(n, n, device=dev, dtype=torch.float64)

    def adj():
        assert la.dgesv_adjoint_async(n, nrhs, lu, n, ipiv, x, n, xbar, n, bbar, n, abar, n) == 0
    out["adjoint_ms"] = timed(adj)
    return out


def mp_compare(n, seed, half=False):
    """Time-to-FP64-backward-error at order n (A, b ~ N(0,1), nrhs = 1): the plain
    FP64 DGESV path against the mixed-precision solvers (DSGESV: FP32 LU + FP64
    refinement; DGESV_IR with an FP32 LU; with half=True also DGESV_IR with
    FP16 / BF16 trailing updates after DGEEQUB, FP64 refinement and GMRES-IR,
    which on N(0,1) at this order needs hundreds of GMRES steps).  Each variant reports its time, ITER, its normwise
    backward error max|b - A x| / (||A||_inf max|x|) computed in FP64 on the GPU,
    and its forward error against the FP64 path's solution."""
    import numpy as np
    import torch
    import lapack_amd as la
    dev = torch.device("cuda:0")
    print(f"[bench] mixed precision: generating A {n}x{n} N(0,1) seed {seed}", file=sys.stderr, flush=True)
    a = la.matrix(3, seed, n, n)
    b = la.matrix(3, seed + 1, n, 1)
    d_a0 = torch.from_numpy(a.T).to(dev)  # row-major view of the column-major bytes (A^T as a tensor)
    del a
    d_b = torch.from_numpy(b.T.copy()).to(dev)
    rowsum = d_a0.abs().sum(dim=0)  # tensor is A^T: column sums of A^T = row sums of A
    anorm = rowsum.max().item()

    def berr(x):
        r = d_b - x @ d_a0  # (1 x n): b^T - x^T A^T
        return (r.abs().max() / (anorm * x.abs().max())).item()

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, (time.perf_counter() - t) * 1e3

    res = {}
    d_a = torch.empty_like(d_a0)
    ipiv = torch.zeros(n, dtype=torch.int32, device=dev)
    # plain FP64 DGETRF + DGETRS (best of 2)
    best = None
    for _ in range(2):
        d_a.copy_(d_a0)
        x64 = d_b.clone()
        info, ms = timed(lambda: la.dgesv(n, 1, d_a, n, ipiv, x64, n))
        assert info == 0, info
        best = ms if best is None else min(best, ms)
    res["fp64_dgesv"] = {"ms": round(best, 2), "berr": berr(x64)}
    xn = x64.abs().max().item()
    variants = [("fp32_dsgesv", None), ("fp32_ir", "S")]
    if half:
        variants += [("fp16_ir", "H"), ("bf16_ir", "B")]
    for name, prec in variants:
        best, out = None, None
        for _ in range(2 if prec in (None, "S") else 1):
            d_a.copy_(d_a0)
            x = torch.zeros_like(d_b)
            if prec is None:
                (info, it), ms = timed(lambda: la.dsgesv(n, 1, d_a, n, ipiv, d_b, n, x, n))
                git = 0
            else:
                (info, it, git), ms = timed(lambda: la.dgesv_ir(n, 1, d_a, n, ipiv, d_b, n, x, n, prec))
            assert info == 0, (name, info)
            if best is None or ms < best:
                best, out = ms, (it, git, x.clone())
        it, git, x = out
        res[name] = {"ms": round(best, 2), "ratio_to_fp64": round(best / res["fp64_dgesv"]["ms"], 4), "iter": it,
                     "gmres_iter": git, "berr": berr(x), "fwd_err_vs_fp64": ((x - x64).abs().max().item() / xn)}
        print(f"[bench] {name}: {res[name]}", file=sys.stderr, flush=True)
    return {"n": n, "nrhs": 1, "data": f"A, b ~ N(0,1) seeds {seed},{seed + 1}", **res}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(this LU path runs on one GPU; multi-GPU sharding is not part of it).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import lapack_amd as la

    n, nrhs, K, W = args.n, args.nrhs, args.steps, args.warmup
    dev = torch.device("cuda:0")
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    print(f"[bench] generating A {n}x{n}, B {n}x{nrhs} (N(0,1), seed {args.seed})", file=sys.stderr, flush=True)
    a = la.matrix(3, args.seed, n, n)
    b = la.matrix(3, args.seed + 1, n, nrhs)
    d_a0 = torch.from_numpy(a.T).to(dev)  # row-major (n, n) view of the column-major bytes
    d_b0 = torch.from_numpy(b.T.copy()).to(dev)
    del a
    nst = K + W
    d_as = [torch.empty_like(d_a0) for _ in range(nst)]
    d_bs = [torch.empty_like(d_b0) for _ in range(nst)]
    for t in d_as:
        t.copy_(d_a0)
    for t in d_bs:
        t.copy_(d_b0)
    d_ipiv = [torch.zeros(n, dtype=torch.int32, device=dev) for _ in range(nst)]
    d_info = torch.zeros(nst, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(nst)]

    def step(i):
        evs[i][0].record(stream)
        rc = la.dgetrf_async(n, n, d_as[i], n, d_ipiv[i], d_info[i:i + 1], stream)
        assert rc == 0, rc
        evs[i][1].record(stream)
        rc = la.dgetrs_async("N", n, nrhs, d_as[i], n, d_ipiv[i], d_bs[i], n, stream)
        assert rc == 0, rc
        evs[i][2].record(stream)

    if args.calib_copy_mib > 0:
        src = torch.empty(args.calib_copy_mib << 17, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for i in range(W):
        step(K + i)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # timed steps: events only around the trailing-update launches (the
    # roofline's kernel); the full phase breakdown comes from one more,
    # untimed step afterwards
    la.phase_reset()
    la.phase_enable(not args.no_phase, level=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(K):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    la.phase_enable(False)
    ph_upd = la.phase_read()
    ph = ph_upd
    if not args.no_phase:
        d_as[0].copy_(d_a0)
        d_bs[0].copy_(d_b0)
        la.phase_reset()
        la.phase_enable(True, level=1)
        step(0)
        torch.cuda.synchronize()
        la.phase_enable(False)
        ph = la.phase_read()
    infos = d_info.cpu().numpy()
    assert (infos == 0).all(), infos

    T = t1 - t0
    f_getrf = 2.0 / 3.0 * n ** 3
    f_getrs = 2.0 * n * n * nrhs
    value = K * (f_getrf + f_getrs) / T / 1e12
    print(f"[bench] {K} steps in {T:.3f} s", file=sys.stderr, flush=True)

    # Backward-error check of the last timed solve: ||B - A X||_1 / (||A||_1 ||X||_1 eps)
    x = d_bs[K - 1].double()
    a0 = d_a0.double()  # row-major view: element (i, j) of A is a0[j, i]
    r = d_b0 - (x @ a0)  # rows of r are columns: r^T = B^T - X^T A^T
    anorm = a0.abs().sum(dim=1).max().item()  # ||A||_1 = max column sum
    ratio = (r.abs().sum(dim=1) / (anorm * x.abs().sum(dim=1) * 2.0 ** -53)).max().item()

    gemm_ms, gemm_calls, gemm_flops = ph_upd["gemm"]
    # DGETRF and DGETRS times of the timed steps (events on the caller's stream)
    getrf_ms = sum(evs[i][0].elapsed_time(evs[i][1]) for i in range(K))
    getrs_ms = sum(evs[i][1].elapsed_time(evs[i][2]) for i in range(K))
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(T / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"DGETRF+DGETRS FP64 n={n} nrhs={nrhs}, A,B ~ N(0,1) seed {args.seed}, device-resident",
            "n": n, "nrhs": nrhs, "nb": la.get_tuning()[0], "leaf": la.get_tuning()[1],
        },
        "roofline": {
            "bound": "mfma",
            "kernel": ("DGETRF trailing update A22 -= A21*U12: update_big_kernel<double,4,2> (256x128 tiles), "
                       "gemm64_kernel<double,true> (64x64 tiles) for the next panel's columns, "
                       "gemm_kernel<double,N,N,update> (128x128) for the last updates beside the panel"),
            "achieved": round(achieved, 3),
            "peak": PEAK_FP64,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64, 4),
            "traffic": update_traffic(n) if nrhs == 16 else None,
            "launches": gemm_calls,
            "avg_launch_ms": round(gemm_ms / max(gemm_calls, 1), 4),
        },
        "getrf_tflops": round(K * f_getrf / (getrf_ms * 1e-3) / 1e12, 3) if getrf_ms > 0 else None,
        "getrf_frac_of_fp64_peak": round(K * f_getrf / (getrf_ms * 1e-3) / 1e12 / PEAK_FP64, 4) if getrf_ms > 0 else None,
        "getrf_ms_per_step": round(getrf_ms / K, 3),
        "getrs_ms_per_step": round(getrs_ms / K, 3),
        "phase_ms_one_step": {p: round(ph[p][0], 3) for p in ph},
        "dget02_ratio": round(ratio, 4),
    }
    if not args.no_mp:
        del d_as, d_bs, d_a0
        torch.cuda.empty_cache()
        out["mp_ir"] = mp_compare(args.mp_n, 32768, args.mp_half)
    if args.extras:
        torch.cuda.empty_cache()
        print(f"[bench] DGEEQU/DGEEQUB/DLAQGE/DGERFS/adjoint at n={n}", file=sys.stderr, flush=True)
        out["extras"] = extras(n, nrhs, args.seed)
    if not args.no_cpu:
        print(f"[bench] CPU baseline sample n={args.cpu_n}", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
