"""GPU parity of the batched uneven LU (one persistent launch per batch).

The reference has no batched routine: the oracle is the per-matrix loop of
reference DGETRF / DGETRS.  Bars per matrix, as for the single-matrix path:
IPIV identical (up to a documented near-tie), INFO identical, DGET01 and
DGET02 ratios < 30 (for FP32 with eps = 2^-24).  The batch workload itself
(10000 systems, n uniform in [16, 512], nrhs in [1, 4], entries
uniform(-1, 1)) runs through an autotuned plan, captured in a hipGraph and
replayed; every replay must reproduce the eager result bit for bit.
"""
import numpy as np
import pytest
import torch

import lapack_amd as la
import oracle

from test_gpu_lu import assert_ipiv

pytestmark = pytest.mark.gpu

THRESH = 30.0


class Batch:
    """Matrices packed back to back in one device buffer, column-major, with
    per-matrix lda >= n; pointer arrays on the device."""

    def __init__(self, ns, nrhs, seed, dtype=np.float64, lda_pad=None):
        self.ns = list(ns)
        self.nrhs = list(nrhs)
        rng = np.random.default_rng(seed)
        self.ldas = [n + (int(rng.integers(0, 3)) if lda_pad else 0) for n in self.ns]
        sizes = [max(1, l * n) for l, n in zip(self.ldas, self.ns)]
        bsizes = [max(1, n * r) for n, r in zip(self.ns, self.nrhs)]
        self.aoff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.boff = np.concatenate([[0], np.cumsum(bsizes)]).astype(np.int64)
        vals = np.empty(int(self.aoff[-1]), dtype=np.float64)
        la.lib().lapack_amd_dlarnv(2, seed, 0, len(vals), vals.ctypes.data)
        bvals = np.empty(int(self.boff[-1]), dtype=np.float64)
        la.lib().lapack_amd_dlarnv(2, seed + 7, 0, len(bvals), bvals.ctypes.data)
        self.dtype = dtype
        self.h_a = vals.astype(dtype)
        self.h_b = bvals.astype(dtype)
        tdt = torch.float64 if dtype == np.float64 else torch.float32
        self.d_a = torch.from_numpy(self.h_a).cuda()
        self.d_a0 = self.d_a.clone()
        self.d_b = torch.from_numpy(self.h_b).cuda()
        self.d_b0 = self.d_b.clone()
        es = self.d_a.element_size()
        nmax = max(self.ns) if self.ns else 1
        self.d_ipiv = torch.zeros(len(self.ns) * max(nmax, 1), dtype=torch.int32, device="cuda")
        self.d_aptr = torch.tensor([self.d_a.data_ptr() + int(o) * es for o in self.aoff[:-1]], dtype=torch.int64,
                                   device="cuda")
        self.d_bptr = torch.tensor([self.d_b.data_ptr() + int(o) * es for o in self.boff[:-1]], dtype=torch.int64,
                                   device="cuda")
        self.d_pptr = torch.tensor([self.d_ipiv.data_ptr() + 4 * k * nmax for k in range(len(self.ns))],
                                   dtype=torch.int64, device="cuda")
        self.d_lda = torch.tensor(self.ldas, dtype=torch.int32, device="cuda")
        self.d_ldb = torch.tensor([max(1, n) for n in self.ns], dtype=torch.int32, device="cuda")
        self.d_nrhs = torch.tensor(self.nrhs, dtype=torch.int32, device="cuda")
        self.d_info = torch.zeros(len(self.ns), dtype=torch.int32, device="cuda")
        self.nmax = nmax

    def a_host(self, k, arr=None):
        arr = self.h_a if arr is None else arr
        n, l = self.ns[k], self.ldas[k]
        blk = arr[self.aoff[k]:self.aoff[k] + l * n]
        return np.asfortranarray(blk.reshape(n, l).T[:n, :n]).astype(np.float64)

    def b_host(self, k, arr=None):
        arr = self.h_b if arr is None else arr
        n, r = self.ns[k], self.nrhs[k]
        return np.asfortranarray(arr[self.boff[k]:self.boff[k] + n * r].reshape(r, n).T).astype(np.float64)


def eps_of(dtype):
    return 2.0 ** -53 if dtype == np.float64 else 2.0 ** -24


def check_matrix(bt, k, lu_all, piv_all, x_all, info_all):
    a = bt.a_host(k)
    n = bt.ns[k]
    lu = bt.a_host(k, lu_all)
    p = piv_all[k * bt.nmax:k * bt.nmax + n]
    scale = 2.0 ** -53 / eps_of(bt.dtype)  # ratios of the oracle are in units of 2^-53
    if bt.dtype == np.float64:
        lur, pr, infor = oracle.dgetrf(a)
    else:
        lur, pr, infor = oracle.sgetrf(a.astype(np.float32))
        lur = lur.astype(np.float64)
    assert info_all[k] == infor, (k, n, info_all[k], infor)
    assert_ipiv(p, pr, lur, unit=eps_of(bt.dtype))
    assert oracle.dget01(a, lu, p) * scale < THRESH, (k, n)
    if infor == 0 and bt.nrhs[k] > 0:
        b = bt.b_host(k)
        x = bt.b_host(k, x_all)
        assert oracle.dget02(a, x, b) * scale < THRESH, (k, n)


def run(bt, plan):
    assert plan.getrf(bt.d_aptr, bt.d_lda, bt.d_pptr, bt.d_info) == 0
    assert plan.getrs(bt.d_nrhs, bt.d_aptr, bt.d_lda, bt.d_pptr, bt.d_bptr, bt.d_ldb) == 0
    torch.cuda.synchronize()
    return bt.d_a.cpu().numpy(), bt.d_ipiv.cpu().numpy(), bt.d_b.cpu().numpy(), bt.d_info.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_uneven_parity(dev, dtype):
    rng = np.random.default_rng(11)
    edge = [1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512]
    ns = edge + [int(v) for v in rng.integers(1, 513, 37)]
    nrhs = [int(v) for v in rng.integers(1, 7, len(ns))]  # > 4 exercises the second pass over B
    bt = Batch(ns, nrhs, 12, dtype, lda_pad=True)
    plan = la.BatchPlan(ns, "d" if dtype == np.float64 else "s", autotune=False)
    lu, piv, x, info = run(bt, plan)
    for k in range(len(ns)):
        check_matrix(bt, k, lu, piv, x, info)


def test_batched_every_config_and_singular(dev):
    rng = np.random.default_rng(21)
    ns = [int(v) for v in rng.integers(16, 300, 24)]
    nrhs = [int(v) for v in rng.integers(1, 5, len(ns))]
    bt = Batch(ns, nrhs, 22)
    # exact zero columns -> INFO = first zero pivot (DCHKGE types 5-7)
    for k, col in ((0, 0), (3, 5), (7, ns[7] - 1)):
        n, l = ns[k], bt.ldas[k]
        bt.h_a[bt.aoff[k] + col * l:bt.aoff[k] + col * l + n] = 0.0
    bt.d_a.copy_(torch.from_numpy(bt.h_a))
    bt.d_a0.copy_(bt.d_a)
    plan = la.BatchPlan(ns, "d", autotune=False)
    results = []
    for nb, nt in ((8, 256), (16, 256), (16, 512), (32, 256), (16, 128), (8, 128)):
        assert plan.set_config(nb, nt) == 0
        bt.d_a.copy_(bt.d_a0)
        bt.d_b.copy_(bt.d_b0)
        lu, piv, x, info = run(bt, plan)
        for k in range(len(ns)):
            n = ns[k]
            a = bt.a_host(k)
            lur, pr, infor = oracle.dgetrf(a)
            assert info[k] == infor
            assert_ipiv(piv[k * bt.nmax:k * bt.nmax + n], pr, lur)
            assert oracle.dget01(a, bt.a_host(k, lu), piv[k * bt.nmax:k * bt.nmax + n]) < THRESH
        results.append(info.copy())
    assert all(np.array_equal(r, results[0]) for r in results)


def test_batched_10000_graph_replay(dev):
    """The batch workload: 10000 systems, n in [16, 512], nrhs in [1, 4],
    autotuned plan, one persistent getrf launch + one getrs launch captured
    in a hipGraph; three replays (inputs restored eagerly in between) equal
    the eager run bitwise; 150 systems checked against the oracle."""
    rng = np.random.default_rng(4)
    ns = [int(v) for v in rng.integers(16, 513, 10000)]
    nrhs = [int(v) for v in rng.integers(1, 5, 10000)]
    bt = Batch(ns, nrhs, 5)
    plan = la.BatchPlan(ns, "d", autotune=True)
    nb, nt, ms = plan.config()
    print(f"autotuned: NB={nb} threads={nt} ({ms:.2f} ms on the tuning sample)")
    lu, piv, x, info = run(bt, plan)
    assert (info == 0).all()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    bt.d_a.copy_(bt.d_a0)
    bt.d_b.copy_(bt.d_b0)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        assert plan.getrf(bt.d_aptr, bt.d_lda, bt.d_pptr, bt.d_info, torch.cuda.current_stream()) == 0
        assert plan.getrs(bt.d_nrhs, bt.d_aptr, bt.d_lda, bt.d_pptr, bt.d_bptr, bt.d_ldb,
                          torch.cuda.current_stream()) == 0
    for rep in range(3):
        bt.d_a.copy_(bt.d_a0)
        bt.d_b.copy_(bt.d_b0)
        bt.d_ipiv.zero_()
        bt.d_info.fill_(-7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(bt.d_a.cpu().numpy().view(np.uint64), lu.view(np.uint64)), rep
        assert np.array_equal(bt.d_ipiv.cpu().numpy(), piv), rep
        assert np.array_equal(bt.d_b.cpu().numpy().view(np.uint64), x.view(np.uint64)), rep
        assert np.array_equal(bt.d_info.cpu().numpy(), info), rep
    for k in sorted(rng.choice(len(ns), 150, replace=False).tolist()):
        check_matrix(bt, k, lu, piv, x, info)
