"""The environment-selectable forms of the LU path, each against the oracle
(IPIV identical or a traced near-tie, DGET01 / DGET02 below 30) on the same
inputs as the default forms:

  * LAPACK_AMD_UPD_PIPE=1: large trailing updates on the persistent,
    software-pipelined 128 x 128 kernel (edge strips on the 128 x 128
    one-tile kernel) instead of the 256 x 128 one-tile kernel;
  * LAPACK_AMD_PANEL_XCD=0: the multi-workgroup panel shares every XCD with
    the update instead of owning XCD 0;
  * LAPACK_AMD_UPD_DMA=0: large trailing updates on the register-staged
    256 x 128 kernel instead of the direct-to-LDS one;
  * LAPACK_AMD_UPD_BIG_MIN=1: every trailing update that fits on the
    direct-to-LDS kernel (persistent beside the panel), at orders where the
    default threshold would leave them on the 128 x 128 kernels.
The FP32 forms of the last two (and stages of 16 k instead of 32 for the
FP32 direct-to-LDS kernel beside the panel) are run by
test_sgetrf_variants_against_reference.
"""
import numpy as np
import pytest

import lapack_amd as la
import oracle

from gpu_util import from_dev, ivec, to_dev
from test_gpu_lu import THRESH, assert_ipiv

pytestmark = pytest.mark.gpu


def _solve(a, b):
    n, nrhs = b.shape
    d_a, d_b = to_dev(a), to_dev(b)
    d_p = ivec(n)
    assert la.dgetrf(n, n, d_a, n, d_p) == 0
    lu = from_dev(d_a, (n, n))
    p = d_p.cpu().numpy()[:n]
    x = {}
    for trans in "NT":
        d_x = to_dev(b)
        assert la.dgetrs(trans, n, nrhs, d_a, n, d_p, d_x, n) == 0
        x[trans] = from_dev(d_x, (n, nrhs))
    return lu, p, x


@pytest.mark.parametrize("env", [{"LAPACK_AMD_UPD_PIPE": "1"}, {"LAPACK_AMD_PANEL_XCD": "0"},
                                 {"LAPACK_AMD_UPD_DMA": "0"}, {"LAPACK_AMD_UPD_BIG_MIN": "1"}],
                         ids=["upd_pipe", "panel_shared", "upd_staged", "upd_dma_all"])
@pytest.mark.parametrize("n,nrhs", [(3000, 16), (4100, 5)])
def test_variant_against_oracle(dev, monkeypatch, env, n, nrhs):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    a = la.matrix(3, 9000 + n, n, n)
    b = la.matrix(3, 9001 + n, n, nrhs)
    lu_r, p_r, info_r = oracle.dgetrf(a)
    assert info_r == 0
    lu, p, x = _solve(a, b)
    assert_ipiv(p, p_r, lu_r, lu)
    assert oracle.dget01(a, lu, p) < THRESH
    for trans in "NT":
        assert oracle.dget02(a, x[trans], b, trans) < THRESH, trans
    assert la.device_status(clear=True) == 0


@pytest.mark.parametrize("env", [{"LAPACK_AMD_UPD_DMA": "0"}, {"LAPACK_AMD_UPD_BIG_MIN": "1"},
                                 {"LAPACK_AMD_UPD_BIG_MIN": "1", "LAPACK_AMD_UPD_DMA_KS": "16"}],
                         ids=["upd_staged", "upd_dma_all", "upd_dma_all_k16"])
@pytest.mark.parametrize("n", [1500, 2048])
def test_sgetrf_variants_against_reference(dev, monkeypatch, env, n):
    """SGETRF under the update variants: IPIV against the reference SGETRF (near-tie bar in units of 2^-24)
    and the SGET01-style ratio, as test_sgetrf_parity."""
    from test_gpu_lu import test_sgetrf_parity
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    test_sgetrf_parity(dev, n)
