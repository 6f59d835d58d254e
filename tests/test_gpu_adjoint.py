"""GPU checks of the solve's adjoint: Bbar = A^-T Xbar, Abar = -Bbar X^T.

Workload: systems A = N(0,1) + 4 sqrt(n) I of order n = 1024, B n x 8
N(0,1), loss = sum(W .* X) with W N(0,1), so that Xbar = W.  Three bars:
  * against the gradient built on the CPU from the reference's DGETRS('T')
    (the flang build in oracle/_ref where present, else its restatement,
    pinned bitwise to it) on the reference DGETRF factors: relative 1e-11;
  * against central finite differences of the loss, h = 1e-6, on 64 seeded
    entries of A and 64 of B: |fd - g| <= 1e-8 + 1e-5 |g|;
  * the rank-NRHS kernel against the oracle's DGEMM('N','T') to DBLAT3's
    componentwise ratio < 16, with BETA = 0 never reading C.
"""
import numpy as np
import pytest
import torch

import lapack_amd as la
import oracle
from lapack_amd.autograd import solve

from gpu_util import from_dev, ivec, to_dev

pytestmark = pytest.mark.gpu

N, NRHS = 1024, 8


def system(seed, n=N, nrhs=NRHS):
    a = la.matrix(3, seed, n, n) + 4.0 * np.sqrt(n) * np.eye(n)
    b = la.matrix(3, seed + 100000, n, nrhs)
    w = la.matrix(3, seed + 200000, n, nrhs)
    return np.asfortranarray(a), b, w


def t(m):
    return torch.from_numpy(np.ascontiguousarray(m)).cuda()


def cpu_gradient(a, b, w):
    """Abar, Bbar of loss = sum(W .* A^-1 B) from the reference's factors and DGETRS('T')."""
    getrf = oracle.ref_dgetrf if oracle.ref_available() else oracle.dgetrf
    getrs = oracle.ref_dgetrs if oracle.ref_available() else oracle.dgetrs
    lu, p, info = getrf(a)[:3]
    assert info == 0
    x, _ = getrs(lu, p, b)
    bbar, _ = getrs(lu, p, w, trans="T")
    abar = -bbar @ x.T
    return abar, bbar


def rel(x, y):
    return np.max(np.abs(x - y)) / np.max(np.abs(y))


def test_adjoint_matches_reference_gradient(dev):
    a, b, w = system(3100)
    A = t(a).requires_grad_()
    B = t(b).requires_grad_()
    X = solve(A, B)
    (t(w) * X).sum().backward()
    abar, bbar = cpu_gradient(a, b, w)
    assert rel(B.grad.cpu().numpy(), bbar) < 1e-11
    assert rel(A.grad.cpu().numpy(), abar) < 1e-11


def loss_gpu(a, b, w):
    n, nrhs = b.shape
    d_a = to_dev(a)
    d_b = to_dev(b)
    info = la.dgesv(n, nrhs, d_a, n, ivec(n), d_b, n)
    assert info == 0
    return float(np.sum(w * from_dev(d_b, (n, nrhs))))


def test_adjoint_central_differences(dev):
    a, b, w = system(3200)
    A = t(a).requires_grad_()
    B = t(b).requires_grad_()
    (t(w) * solve(A, B)).sum().backward()
    ga = A.grad.cpu().numpy()
    gb = B.grad.cpu().numpy()
    rng = np.random.default_rng(3201)
    h = 1e-6
    for which, g, m in (("A", ga, a), ("B", gb, b)):
        idx = [(int(rng.integers(m.shape[0])), int(rng.integers(m.shape[1]))) for _ in range(64)]
        for (i, j) in idx:
            ap, am, bp, bm = a.copy(order="F"), a.copy(order="F"), b.copy(order="F"), b.copy(order="F")
            if which == "A":
                ap[i, j] += h
                am[i, j] -= h
            else:
                bp[i, j] += h
                bm[i, j] -= h
            fd = (loss_gpu(ap, bp, w) - loss_gpu(am, bm, w)) / (2 * h)
            assert abs(fd - g[i, j]) <= 1e-8 + 1e-5 * abs(g[i, j]), (which, i, j, fd, g[i, j])


def test_adjoint_batch_256(dev):
    """The batched workload: 256 systems of order 1024 through one autograd
    graph; four of them checked against the CPU gradient, all for finite
    values and the identity Abar = -Bbar X^T."""
    batch = 256
    rng = np.random.default_rng(3300)
    check = sorted(rng.choice(batch, 4, replace=False).tolist())
    A = torch.empty(batch, N, N, dtype=torch.float64, device="cuda")
    B = torch.empty(batch, N, NRHS, dtype=torch.float64, device="cuda")
    W = torch.empty(batch, N, NRHS, dtype=torch.float64, device="cuda")
    host = {}
    for k in range(batch):
        a, b, w = system(3300 + k)
        if k in check:
            host[k] = (a, b, w)
        A[k].copy_(t(a))
        B[k].copy_(t(b))
        W[k].copy_(t(w))
    A.requires_grad_()
    B.requires_grad_()
    X = solve(A, B)
    (W * X).sum().backward()
    assert torch.isfinite(A.grad).all() and torch.isfinite(B.grad).all()
    for k in check:
        abar, bbar = cpu_gradient(*host[k])
        assert rel(B.grad[k].cpu().numpy(), bbar) < 1e-11, k
        assert rel(A.grad[k].cpu().numpy(), abar) < 1e-11, k
    k = check[0]
    outer = -(B.grad[k] @ X[k].detach().t())
    assert torch.allclose(A.grad[k], outer, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("m,n,k,beta", [(1, 1, 1, 0.0), (100, 70, 8, 0.0), (257, 300, 3, 1.0), (64, 64, 32, 1.0),
                                        (130, 90, 40, 1.0)])
def test_rankk_kernel_against_dgemm(dev, m, n, k, beta):
    rng = np.random.default_rng(m * 1000 + k)
    bb = np.asfortranarray(rng.standard_normal((m, k)))
    x = np.asfortranarray(rng.standard_normal((n, k)))
    c0 = np.asfortranarray(rng.standard_normal((m, n)))
    if beta == 0.0:
        c0[0, 0] = np.nan  # BETA = 0: C is not read (BLAS/SRC/dgemm.f:256-262)
    # the adjoint entry point with identity factors computes Abar = -Xbar X^T (+ C)
    # on an n x n system; here the rank-k product is checked directly through DGEMM's
    # definition via a square embedding.
    s = max(m, n)
    lu = np.asfortranarray(np.eye(s))
    xs = np.zeros((s, k), order="F")
    xs[:n] = x
    xb = np.zeros((s, k), order="F")
    xb[:m] = bb
    cs = np.zeros((s, s), order="F")
    cs[:m, :n] = c0
    d_lu, d_p = to_dev(lu), torch.arange(1, s + 1, dtype=torch.int32, device="cuda")
    d_x, d_xb, d_bb, d_c = to_dev(xs), to_dev(xb), to_dev(np.zeros((s, k))), to_dev(cs)
    rc = la.dgesv_adjoint_async(s, k, d_lu, s, d_p, d_x, s, d_xb, s, d_bb, s, d_c, s, beta != 0.0)
    torch.cuda.synchronize()
    assert rc == 0
    got = from_dev(d_c, (s, s))[:m, :n]
    want = oracle.dgemm("N", "T", -1.0, bb, x, 1.0 if beta != 0.0 else 0.0, np.nan_to_num(c0))
    eps = 2.0 ** -52
    g = np.abs(bb) @ np.abs(x).T + (np.abs(c0) if beta != 0.0 else 0.0)
    ratio = np.max(np.abs(got - want) / (eps * np.maximum(g, 1e-300)))
    assert np.isfinite(got).all() and ratio < 16
    # Bbar of identity factors is Xbar itself
    assert np.array_equal(from_dev(d_bb, (s, k)), xb)


def _pack(mats, dtype=torch.float64):
    """Column-major matrices packed in one device buffer: (buffer, int64
    pointer tensor, int32 leading dimensions)."""
    offs = np.cumsum([0] + [m.size for m in mats])
    buf = torch.zeros(int(offs[-1]), dtype=dtype, device="cuda")
    for m, o in zip(mats, offs[:-1]):
        buf[int(o):int(o) + m.size] = torch.from_numpy(np.ravel(np.asfortranarray(m), order="F").copy()).cuda()
    es = buf.element_size()
    ptrs = torch.tensor([buf.data_ptr() + int(o) * es for o in offs[:-1]], dtype=torch.int64, device="cuda")
    lds = torch.tensor([m.shape[0] for m in mats], dtype=torch.int32, device="cuda")
    return buf, ptrs, lds, offs


def _unpack(buf, offs, shapes):
    h = buf.cpu().numpy()
    return [np.asfortranarray(h[int(o):int(o) + r * c].reshape(c, r).T) for o, (r, c) in zip(offs[:-1], shapes)]


def test_batched_adjoint_matches_reference_gradient(dev):
    """The batched adjoint (one transposed-solve launch and one rank-NRHS
    launch over every system) on the batched factors, uneven orders and
    right-hand-side counts: Bbar and Abar against the reference's gradient
    (DGETRS('T') on the reference's factors), relative 1e-11, and against the
    single-system entry point; accumulate adds to Abar."""
    ns = [1024, 17, 64, 65, 300, 513, 1024, 700, 1, 128]
    rs = [8, 1, 3, 4, 8, 5, 2, 7, 1, 33]
    systems = [system(4000 + k, n, r) for k, (n, r) in enumerate(zip(ns, rs))]
    plan = la.BatchPlan(ns, "d", autotune=False)
    bufa, pa, lda, offa = _pack([s[0] for s in systems])
    bufb, pb, ldb, offb = _pack([s[1] for s in systems])
    bufw, pw, ldw, offw = _pack([s[2] for s in systems])
    bufab, pab, ldab, offab = _pack([np.full((n, n), 0.5) for n in ns])
    piv = torch.zeros(int(sum(ns)), dtype=torch.int32, device="cuda")
    pofs = np.cumsum([0] + ns)
    pp = torch.tensor([piv.data_ptr() + 4 * int(o) for o in pofs[:-1]], dtype=torch.int64, device="cuda")
    info = torch.zeros(len(ns), dtype=torch.int32, device="cuda")
    nrhs = torch.tensor(rs, dtype=torch.int32, device="cuda")
    assert plan.getrf(pa, lda, pp, info) == 0
    assert plan.getrs(nrhs, pa, lda, pp, pb, ldb) == 0          # X in place of B
    assert plan.adjoint(nrhs, pa, lda, pp, pb, ldb, pw, ldw, pab, ldab, accumulate=True) == 0
    torch.cuda.synchronize()
    assert int(info.abs().sum()) == 0
    bbars = _unpack(bufw, offw, [(n, r) for n, r in zip(ns, rs)])
    abars = _unpack(bufab, offab, [(n, n) for n in ns])
    for k, (a, b, w) in enumerate(systems):
        abar_r, bbar_r = cpu_gradient(a, b, w)
        assert rel(bbars[k], bbar_r) < 1e-11, k
        assert rel(abars[k] - 0.5, abar_r) < 1e-11, k
    print(f"batched adjoint: {len(ns)} systems, orders {ns}, nrhs {rs}: Bbar, Abar within 1e-11 of the reference")
