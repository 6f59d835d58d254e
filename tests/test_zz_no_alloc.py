"""GPU test of lapack_amd_dgetrs_noalloc_async, the DGETRS entry point that
allocates nothing (its workspace is the library's static device array):
through the C-ABI, against lapack_amd_dgetrs_async on the same factors and
right-hand sides -- the same kernels, so bitwise equal -- on two calls in a
row, the second smaller than the first (the array keeps the first call's
contents), for 'N' and 'T'; and the size limit."""
import numpy as np
import pytest

import lapack_amd as la

from gpu_util import from_dev, ivec, to_dev
from test_gpu_lu import gpu_getrf

pytestmark = pytest.mark.gpu


def _solve_both(lu, p, b, trans):
    import torch
    n, nrhs = b.shape
    d_lu = to_dev(lu)
    d_p = to_dev(p.reshape(-1, 1).astype(np.int32))
    d_b1, d_b2 = to_dev(b), to_dev(b)
    assert la.dgetrs_async(trans, n, nrhs, d_lu, n, d_p, d_b1, n) == 0
    d_info = ivec(1)
    d_info.fill_(77)
    assert la.dgetrs_noalloc_async(trans, n, nrhs, d_lu, n, d_p, d_b2, n, d_info) == 0
    torch.cuda.synchronize()
    assert int(d_info.cpu()[0]) == 0
    return from_dev(d_b1, (n, nrhs)), from_dev(d_b2, (n, nrhs))


@pytest.mark.parametrize("trans", ["N", "T"])
def test_noalloc_getrs_matches_getrs(dev, trans):
    import oracle
    for n, nrhs, seed in ((3000, 16, 71), (700, 3, 72)):
        a = la.matrix(2, seed, n, n)
        b = la.matrix(2, seed + 1, n, nrhs)
        lu, p, info = gpu_getrf(a)
        assert info == 0
        x_ref, x_na = _solve_both(lu, p, b, trans)
        print(f"noalloc DGETRS('{trans}') n={n} nrhs={nrhs}: bitwise equal to dgetrs_async "
              f"{np.array_equal(x_ref.view(np.int64), x_na.view(np.int64))}", flush=True)
        assert np.array_equal(x_ref.view(np.int64), x_na.view(np.int64))
        assert oracle.dget02(a, x_na, b, trans) < 30


def test_noalloc_limits(dev):
    lib = la.lib()
    cap = lib.lapack_amd_noalloc_scratch_bytes()
    assert cap == 16 << 20
    assert lib.lapack_amd_dgetrs_noalloc_bytes(16384, 16) <= cap
    n, nrhs = 16384, 200
    assert lib.lapack_amd_dgetrs_noalloc_bytes(n, nrhs) > cap
    d_a = to_dev(np.zeros((1, 1)))
    d_info = ivec(1)
    assert la.dgetrs_noalloc_async("N", n, nrhs, d_a, n, ivec(1), d_a, n, d_info) == la.LA_ERR_LIMIT
    # argument checks in DGETRS's order (SRC/dgetrs.f:157-174)
    assert la.dgetrs_noalloc_async("X", 4, 1, d_a, 4, ivec(4), d_a, 4, d_info) == -1
    assert la.dgetrs_noalloc_async("N", 4, 1, d_a, 3, ivec(4), d_a, 4, d_info) == -5
