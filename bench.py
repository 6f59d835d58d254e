#!/usr/bin/env python3
"""bench.py -- DGETRF + DGETRS throughput on one MI355X.

Workload (one "step"): the plain FP64 LU solve at n=16384, nrhs=16, A and B
i.i.d. N(0,1) from the library's seeded generator: DGETRF on A, then
DGETRS('N') on B, both on device-resident inputs through the device entry
points.  Each step works on its own copy of A and B, staged in HBM before
the timed region.

Printed: ONE JSON line with the metric, the GEMM roofline (trailing-update
kernel, FP64 matrix peak) measured with HIP events inside the timed region,
DGETRS and DLASWP bandwidth against HBM, the mixed-precision comparison at
n=32768 (time to the FP64 backward error, FP32 factors against the FP64
path), the other workloads (the n=1024 solve, the 10000-system batch in FP64
and FP32, the 256-system adjoint), and CPU baselines: the reference DGETRF +
reference BLAS built from the reference sources (oracle/_ref) -- or its C++
restatement where that build is absent -- on one host core at n=2048 and
4096 with a labelled n^3 extrapolation to the benchmark's order, and the
batched workload over every host core.
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
PEAK_FP32 = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM = 8000.0  # GB/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--nrhs", type=int, default=16)
    p.add_argument("--seed", type=int, default=16384)
    p.add_argument("--cpu-n", default="2048,4096", help="CPU baseline sample orders (comma-separated)")
    p.add_argument("--no-tune", action="store_true", help="skip the NB / leaf autotuner (library defaults)")
    p.add_argument("--no-workloads", action="store_true",
                   help="skip the n=1024, batched and adjoint workloads (and their CPU baseline)")
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


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def cpu_baseline(sizes, seed, n_target):
    """Reference DGETRF (flang build of the reference sources, 1 core) on N(0,1)
    matrices of the sample orders; the benchmark's order is reached by an n^3
    extrapolation from the largest sample, labelled as such."""
    import lapack_amd as la
    import oracle
    kind = "reference" if oracle.ref_available() else "port"
    fn = oracle.ref_dgetrf if kind == "reference" else oracle.dgetrf
    times = {}
    for n_s in sizes:
        a = la.matrix(3, seed, n_s, n_s)
        t = time.perf_counter()
        _, _, info = fn(a)
        times[n_s] = time.perf_counter() - t
        assert info == 0
    n_big = max(sizes)
    rate = 2.0 / 3.0 * n_big ** 3 / times[n_big]
    est = 2.0 / 3.0 * n_target ** 3 / rate
    model, _ = cpu_model()
    return {
        "value": rate / 1e12,
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": kind,
        "cpu": model,
        "sample": ("DGETRF N(0,1), reference BLAS, one core: " +
                   ", ".join(f"n={k} {v:.2f} s ({2.0 / 3.0 * k ** 3 / v / 1e9:.2f} GFLOP/s)" for k, v in times.items())),
        "n_target_seconds_extrapolated": round(est, 1),
        "extrapolation": f"n^3 from the measured n={n_big} rate to n={n_target} (not measured)",
    }


def _c4_cpu_chunk(args):
    """One worker's share of the CPU batch: reference DGETRF + DGETRS('N') per
    system (numpy-generated uniform(-1,1) entries of the same shapes)."""
    import numpy as np
    import oracle
    ns, nrhs, seed, use_ref = args
    rng = np.random.default_rng(seed)
    getrf = oracle.ref_dgetrf if use_ref else oracle.dgetrf
    getrs = oracle.ref_dgetrs if use_ref else oracle.dgetrs
    for n, r in zip(ns, nrhs):
        a = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
        b = np.asfortranarray(rng.uniform(-1, 1, (n, r)))
        lu, ipiv, info = getrf(a)
        getrs(lu, ipiv, b)
    return len(ns)


def cpu_c4_baseline(ns, nrhs, workers):
    """The batched workload on every host core: the systems dealt round-robin
    to `workers` processes, each running reference DGETRF + DGETRS per system."""
    import multiprocessing as mp
    import oracle
    use_ref = oracle.ref_available()
    chunks = [(ns[w::workers], nrhs[w::workers], 100 + w, use_ref) for w in range(workers)]
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        t = time.perf_counter()
        done = sum(pool.map(_c4_cpu_chunk, chunks))
        dt = time.perf_counter() - t
    flops = sum(2.0 / 3.0 * n ** 3 + 2.0 * n * n * r for n, r in zip(ns, nrhs))
    model, _ = cpu_model()
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": workers,
            "kind": "reference" if use_ref else "port", "cpu": model,
            "sample": f"all {done} systems, DGETRF + DGETRS per system over {workers} processes, {dt:.2f} s"}


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


def timed_events(fn, reps, stream):
    """Mean device time (ms) of fn() over reps calls, HIP events on stream."""
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def c1_solve(seed=1988):
    """The reference's own CPU-sized case: DGETRF + DGETRS, n = 1024, one
    right-hand side, entries uniform(-1, 1); device-resident inputs, each
    repetition on a fresh copy (copies outside the timed region)."""
    import torch
    import lapack_amd as la
    n, reps = 1024, 20
    stream = torch.cuda.current_stream()
    a = torch.from_numpy(la.matrix(2, seed, n, n).T.copy()).cuda()
    b = torch.from_numpy(la.matrix(2, seed + 1, n, 1).T.copy()).cuda()
    aa = [a.clone() for _ in range(reps + 1)]
    bb = [b.clone() for _ in range(reps + 1)]
    ip = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    k = [0]

    def step():
        i = k[0]
        k[0] += 1
        la.dgetrf_async(n, n, aa[i], n, ip, info, stream)
        la.dgetrs_async("N", n, 1, aa[i], n, ip, bb[i], n, stream)
    ms = timed_events(step, reps, stream)
    assert int(info.cpu()[0]) == 0
    flops = 2.0 / 3.0 * n ** 3 + 2.0 * n * n
    return {"workload": "DGETRF+DGETRS FP64 n=1024 nrhs=1, uniform(-1,1)", "ms": round(ms, 4),
            "tflops": round(flops / ms / 1e9, 4), "frac_fp64_peak": round(flops / ms / 1e9 / PEAK_FP64, 4)}


def laswp_bw(d_lu, d_ipiv, n, stream):
    """DLASWP (SRC/dlaswp.f) of all n interchanges of a factorization over the
    n columns of its matrix: algorithmic bytes 16 * n * (rows the permutation
    moves), read and written once."""
    import numpy as np
    import lapack_amd as la
    ipiv = d_ipiv.cpu().numpy()
    perm = np.arange(n)
    for k in range(n):
        p = ipiv[k] - 1
        perm[k], perm[p] = perm[p], perm[k]
    moved = int((perm != np.arange(n)).sum())
    ms = timed_events(lambda: la.dlaswp_async(n, d_lu, n, 1, n, d_ipiv, 1, stream), 3, stream)
    byt = 16.0 * n * moved
    return {"ms": round(ms, 3), "rows_moved": moved, "bytes": int(byt), "gbps": round(byt / ms / 1e6, 1),
            "hbm_frac": round(byt / ms / 1e6 / PEAK_HBM, 4)}


def batch_workload(precision, seed=4):
    """The batched workload: 10000 systems, n uniform in [16, 512], nrhs in
    [1, 4], entries uniform(-1, 1) from the library's generator, each matrix
    its own column-major array (lda = n) behind a device pointer array; an
    autotuned plan, DGETRF then DGETRS over the whole batch per step."""
    import numpy as np
    import torch
    import lapack_amd as la
    rng = np.random.default_rng(seed)
    ns = [int(v) for v in rng.integers(16, 513, 10000)]
    nrhs = [int(v) for v in rng.integers(1, 5, 10000)]
    dt = np.float64 if precision == "d" else np.float32
    aoff = np.concatenate([[0], np.cumsum([n * n for n in ns])]).astype(np.int64)
    boff = np.concatenate([[0], np.cumsum([n * r for n, r in zip(ns, nrhs)])]).astype(np.int64)
    va = la.dlarnv(2, seed, int(aoff[-1])).astype(dt)
    vb = la.dlarnv(2, seed + 7, int(boff[-1])).astype(dt)
    d_a0 = torch.from_numpy(va).cuda()
    d_b0 = torch.from_numpy(vb).cuda()
    d_a, d_b = d_a0.clone(), d_b0.clone()
    es = d_a.element_size()
    nmax = max(ns)
    d_ip = torch.zeros(len(ns) * nmax, dtype=torch.int32, device="cuda")
    aptr = torch.tensor([d_a.data_ptr() + int(o) * es for o in aoff[:-1]], dtype=torch.int64, device="cuda")
    bptr = torch.tensor([d_b.data_ptr() + int(o) * es for o in boff[:-1]], dtype=torch.int64, device="cuda")
    pptr = torch.tensor([d_ip.data_ptr() + 4 * k * nmax for k in range(len(ns))], dtype=torch.int64, device="cuda")
    lda = torch.tensor(ns, dtype=torch.int32, device="cuda")
    nr = torch.tensor(nrhs, dtype=torch.int32, device="cuda")
    info = torch.zeros(len(ns), dtype=torch.int32, device="cuda")
    plan = la.BatchPlan(ns, precision, autotune=True)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    reps, t_f, t_s = 3, 0.0, 0.0
    for r in range(reps + 1):
        d_a.copy_(d_a0)
        d_b.copy_(d_b0)
        torch.cuda.synchronize()
        ev[0].record(stream)
        assert plan.getrf(aptr, lda, pptr, info, stream) == 0
        ev[1].record(stream)
        assert plan.getrs(nr, aptr, lda, pptr, bptr, lda, stream) == 0
        ev[2].record(stream)
        torch.cuda.synchronize()
        if r > 0:
            t_f += ev[0].elapsed_time(ev[1]) / reps
            t_s += ev[1].elapsed_time(ev[2]) / reps
    assert int((info != 0).sum()) == 0
    f_f = sum(2.0 / 3.0 * n ** 3 for n in ns)
    f_s = sum(2.0 * n * n * r for n, r in zip(ns, nrhs))
    byt = sum(2.0 * es * n * n for n in ns)
    peak = PEAK_FP64 if precision == "d" else PEAK_FP32
    nb, nt, _ = plan.config()
    out = {"workload": f"10000 uneven systems ({'FP64' if precision == 'd' else 'FP32'}), n in [16,512], nrhs in [1,4]",
           "getrf_ms": round(t_f, 3), "getrs_ms": round(t_s, 3),
           "tflops": round((f_f + f_s) / (t_f + t_s) / 1e9, 3),
           "getrf_tflops": round(f_f / t_f / 1e9, 3),
           "frac_peak": round(f_f / t_f / 1e9 / peak, 4),
           "getrf_hbm_gbps": round(byt / t_f / 1e6, 1), "getrf_hbm_frac": round(byt / t_f / 1e6 / PEAK_HBM, 4),
           "tuned": {"nb": nb, "threads": nt}}
    plan.close()
    return out, ns, nrhs


def c5_adjoint(seed=256):
    """The adjoint workload: 256 systems, n = 1024, A ~ N(0,1) + 4 sqrt(n) I,
    B 1024 x 8 ~ N(0,1), Xbar ~ N(0,1).  Forward: the batched DGETRF and
    DGETRS('N') (one persistent launch each, a workgroup per system; plan
    and autotuning outside the timed region); backward: the batched adjoint,
    Bbar = A^-T Xbar by DGETRS('T') on the saved factors and Abar = -Bbar X^T
    (two launches over all systems).  The per-system entry points are timed
    beside it for the backward pass (256 x lapack_amd_dgesv_adjoint_async)."""
    import torch
    import lapack_amd as la
    nb, n, nrhs = 256, 1024, 8
    stream = torch.cuda.current_stream()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    a0 = torch.randn(nb, n, n, generator=g, device="cuda", dtype=torch.float64)
    a0 += 4.0 * n ** 0.5 * torch.eye(n, device="cuda", dtype=torch.float64)
    b0 = torch.randn(nb, nrhs, n, generator=g, device="cuda", dtype=torch.float64)
    xbar = torch.randn(nb, nrhs, n, generator=g, device="cuda", dtype=torch.float64)
    lu, x = a0.clone(), b0.clone()
    ip = torch.zeros(nb, n, dtype=torch.int32, device="cuda")
    info = torch.zeros(nb, dtype=torch.int32, device="cuda")
    bbar = torch.empty_like(xbar)
    abar = torch.empty_like(a0)

    def ptrs(t):
        step = t[0].numel() * t.element_size()
        return torch.tensor([t.data_ptr() + k * step for k in range(t.shape[0])], dtype=torch.int64, device="cuda")

    ld = torch.full((nb,), n, dtype=torch.int32, device="cuda")
    nr = torch.full((nb,), nrhs, dtype=torch.int32, device="cuda")
    p_lu, p_ip, p_x, p_bb, p_ab = ptrs(lu), ptrs(ip), ptrs(x), ptrs(bbar), ptrs(abar)
    plan = la.BatchPlan([n] * nb, "d", autotune=True)

    def fwd():
        lu.copy_(a0)
        x.copy_(b0)
        assert plan.getrf(p_lu, ld, p_ip, info, stream) == 0
        assert plan.getrs(nr, p_lu, ld, p_ip, p_x, ld, stream) == 0

    def bwd():
        bbar.copy_(xbar)
        assert plan.adjoint(nr, p_lu, ld, p_ip, p_x, ld, p_bb, ld, p_ab, ld, False, stream) == 0

    def bwd_single():
        for k in range(nb):
            la.dgesv_adjoint_async(n, nrhs, lu[k], n, ip[k], x[k], n, xbar[k], n, bbar[k], n, abar[k], n, False,
                                   stream)
    t_copy = timed_events(lambda: (lu.copy_(a0), x.copy_(b0)), 2, stream)
    t_f = timed_events(fwd, 2, stream) - t_copy
    t_bc = timed_events(lambda: bbar.copy_(xbar), 2, stream)
    t_b = timed_events(bwd, 2, stream) - t_bc
    t_b1 = timed_events(bwd_single, 1, stream)
    assert int((info != 0).sum()) == 0
    nbk, ntk, _ = plan.config()
    plan.close()
    f_fwd = nb * (2.0 / 3.0 * n ** 3 + 2.0 * n * n * nrhs)
    f_bwd = nb * (2.0 * n * n * nrhs * 2)
    byt_bwd = nb * (8.0 * n * n + 8.0 * n * n)
    return {"workload": "adjoint of X = A^-1 B: 256 systems n=1024 nrhs=8 (batched launches)",
            "forward_ms": round(t_f, 2), "forward_tflops": round(f_fwd / t_f / 1e9, 3),
            "forward_frac_fp64_peak": round(f_fwd / t_f / 1e9 / PEAK_FP64, 4),
            "forward_tuned": {"nb": nbk, "threads": ntk},
            "backward_ms": round(t_b, 3), "backward_gbps": round(byt_bwd / t_b / 1e6, 1),
            "backward_hbm_frac": round(byt_bwd / t_b / 1e6 / PEAK_HBM, 4), "backward_gflops": round(f_bwd / t_b / 1e6, 1),
            "backward_per_system_calls_ms": round(t_b1, 2)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(this LU path runs on one GPU; multi-GPU sharding is not part of it).", file=sys.stderr)
        sys.exit(2)
    # the all-core CPU baseline of the batched workload runs first, in
    # processes forked before this one touches the GPU
    cpu_c4 = None
    if not args.no_cpu and not args.no_workloads:
        import numpy as np
        rng = np.random.default_rng(4)
        c4_ns = [int(v) for v in rng.integers(16, 513, 10000)]
        c4_nrhs = [int(v) for v in rng.integers(1, 5, 10000)]
        print("[bench] CPU baseline of the batched workload on every host core", file=sys.stderr, flush=True)
        cpu_c4 = cpu_c4_baseline(c4_ns, c4_nrhs, max(1, min(16, cpu_model()[1])))
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
    tuned = None
    if not args.no_tune:
        # NB and leaf width from the library's autotuner at this order (the
        # table is also written to $LAPACK_AMD_ARTIFACT_DIR when that is set)
        out_dir = os.environ.get("LAPACK_AMD_ARTIFACT_DIR", "")
        path = os.path.join(out_dir, "getrf_tune.json") if out_dir and os.path.isdir(out_dir) else None
        print(f"[bench] autotuning NB / leaf at n={n}", file=sys.stderr, flush=True)
        t_nb, t_leaf, t_ms = la.tune_getrf(n, path)
        tuned = {"nb": t_nb, "leaf": t_leaf, "getrf_ms": round(t_ms, 3),
                 "candidates": "NB in {256, 192, 128} x leaf in {64, 32}, best of 2 timed calls each"}
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
    getrs_bytes = 8.0 * n * n + 16.0 * n * nrhs + 4.0 * n
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
            "n": n, "nrhs": nrhs, "nb": la.get_tuning_for(n)[0], "leaf": la.get_tuning_for(n)[1],
            "autotuned": tuned,
        },
        "roofline": {
            "bound": "mfma",
            "kernel": ("DGETRF trailing update A22 -= A21*U12: update_dma_kernel<double> (256x128 tiles, "
                       "direct global-to-LDS loads), "
                       "gemm64_kernel<double,true> (64x64 tiles) for the next panel's columns, "
                       "gemm_kernel<double,N,N,update> (128x128) for the last updates beside the panel"),
            "achieved": round(achieved, 3),
            "peak": PEAK_FP64,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64, 4),
            "traffic": update_traffic(n) if nrhs == 16 else None,
            "traffic_source": ("static: HBM bytes per update launch from the rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                               "passes summarised in profiles/dgetrf_n16384_update_traffic.md, not measured in this run"),
            "launches": gemm_calls,
            "avg_launch_ms": round(gemm_ms / max(gemm_calls, 1), 4),
        },
        "getrf_tflops": round(K * f_getrf / (getrf_ms * 1e-3) / 1e12, 3) if getrf_ms > 0 else None,
        "getrf_frac_of_fp64_peak": round(K * f_getrf / (getrf_ms * 1e-3) / 1e12 / PEAK_FP64, 4) if getrf_ms > 0 else None,
        "getrf_ms_per_step": round(getrf_ms / K, 3),
        "getrs_ms_per_step": round(getrs_ms / K, 3),
        "getrs": {"bytes": int(getrs_bytes), "gbps": round(getrs_bytes / (getrs_ms / K) / 1e6, 1),
                  "hbm_frac": round(getrs_bytes / (getrs_ms / K) / 1e6 / PEAK_HBM, 4),
                  "algorithmic": "8 n^2 (L and U read once) + 16 n nrhs + 4 n"},
        "phase_ms_one_step": {p: round(ph[p][0], 3) for p in ph},
        "dget02_ratio": round(ratio, 4),
    }
    out["dlaswp"] = laswp_bw(d_as[0], d_ipiv[0], n, stream)
    if not args.no_mp:
        del d_as, d_bs, d_a0
        torch.cuda.empty_cache()
        out["mp_ir"] = mp_compare(args.mp_n, 32768, args.mp_half)
    if args.extras:
        torch.cuda.empty_cache()
        print(f"[bench] DGEEQU/DGEEQUB/DLAQGE/DGERFS/adjoint at n={n}", file=sys.stderr, flush=True)
        out["extras"] = extras(n, nrhs, args.seed)
    if not args.no_workloads:
        torch.cuda.empty_cache()
        wl = {}
        print("[bench] n=1024 solve", file=sys.stderr, flush=True)
        wl["c1_n1024"] = c1_solve()
        for prec in ("d", "s"):
            print(f"[bench] batched workload ({prec})", file=sys.stderr, flush=True)
            wl[f"c4_batched_{'fp64' if prec == 'd' else 'fp32'}"] = batch_workload(prec)[0]
            torch.cuda.empty_cache()
        print("[bench] adjoint workload", file=sys.stderr, flush=True)
        wl["c5_adjoint"] = c5_adjoint()
        out["workloads"] = wl
    if not args.no_cpu:
        sizes = [int(v) for v in args.cpu_n.split(",")]
        print(f"[bench] CPU baseline samples n={sizes}", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sizes, args.seed, n)
        if cpu_c4 is not None:
            out["cpu_baseline_c4"] = cpu_c4
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
