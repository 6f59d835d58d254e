"""torch.autograd harness around the library's solve and its adjoint.

`solve(A, B)` returns X = A^-1 B for a square A (n x n) and B (n x nrhs),
or batches of them (leading batch dimension), as torch tensors on a GPU in
ordinary (row-major) layout.  Forward: DGETRF + DGETRS('N') through the
library's device entry points; the LU factors and IPIV are saved.
Backward: the library's adjoint entry point -- Bbar = A^-T Xbar by
DGETRS('T') on the saved factors (no refactorization) and Abar = -Bbar X^T
by its rank-NRHS kernel.

The library works on column-major storage: the row-major memory of a
tensor M read column-major is M^T, so a column-major copy of A is
A.t().contiguous() and results are read back through .t().
"""
import torch

import lapack_amd as la


def _col_major(m):
    """Row-major (n, k) tensor -> contiguous tensor whose memory is m column-major."""
    return m.t().contiguous()


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B):
        if A.dtype != torch.float64 or B.dtype != torch.float64:
            raise TypeError("solve: float64 tensors expected")
        if not A.is_cuda or not B.is_cuda:
            raise ValueError("solve: GPU tensors expected (the library has no CPU path)")
        n = A.shape[-1]
        if A.shape[-2] != n or B.shape[-2] != n:
            raise ValueError(f"solve: shapes {tuple(A.shape)} and {tuple(B.shape)} do not match")
        batched = A.dim() == 3
        As = A if batched else A.unsqueeze(0)
        Bs = B if batched else B.unsqueeze(0)
        nrhs = Bs.shape[-1]
        stream = torch.cuda.current_stream()
        lus, ipivs, xs = [], [], []
        info = torch.zeros(As.shape[0], dtype=torch.int32, device=A.device)
        for k in range(As.shape[0]):
            lu = _col_major(As[k])
            ipiv = torch.zeros(n, dtype=torch.int32, device=A.device)
            rc = la.dgetrf_async(n, n, lu, n, ipiv, info[k:k + 1], stream)
            if rc != 0:
                raise RuntimeError(f"dgetrf_async returned {rc}")
            x = _col_major(Bs[k])
            rc = la.dgetrs_async("N", n, nrhs, lu, n, ipiv, x, n, stream)
            if rc != 0:
                raise RuntimeError(f"dgetrs_async returned {rc}")
            lus.append(lu)
            ipivs.append(ipiv)
            xs.append(x)
        bad = torch.nonzero(info).flatten()
        if len(bad):
            k = int(bad[0])
            raise RuntimeError(f"solve: matrix {k} is singular (INFO = {int(info[k])})")
        ctx.save_for_backward(torch.stack(lus), torch.stack(ipivs), torch.stack(xs))
        ctx.batched = batched
        X = torch.stack([x.view(nrhs, n).t() for x in xs])
        return X if batched else X[0]

    @staticmethod
    def backward(ctx, Xbar):
        lus, ipivs, xs = ctx.saved_tensors
        n = lus.shape[-1]
        nrhs = xs.shape[1]  # each saved X is (nrhs, n) row-major == n x nrhs column-major
        Xb = Xbar if ctx.batched else Xbar.unsqueeze(0)
        stream = torch.cuda.current_stream()
        abar = torch.empty_like(lus)
        bbar = torch.empty((lus.shape[0], nrhs, n), dtype=lus.dtype, device=lus.device)
        for k in range(lus.shape[0]):
            xb = _col_major(Xb[k].contiguous())
            rc = la.dgesv_adjoint_async(n, nrhs, lus[k], n, ipivs[k], xs[k], n, xb, n, bbar[k], n, abar[k], n,
                                        False, stream)
            if rc != 0:
                raise RuntimeError(f"dgesv_adjoint_async returned {rc}")
        gA = abar.transpose(-1, -2)
        gB = bbar.transpose(-1, -2)
        if not ctx.batched:
            gA, gB = gA[0], gB[0]
        return gA, gB


def solve(A, B):
    """X = A^-1 B, differentiable in A and B."""
    return _Solve.apply(A, B)
