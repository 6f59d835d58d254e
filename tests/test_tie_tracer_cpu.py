"""CPU checks of the IPIV near-tie tracer the GPU tests use (assert_ipiv in
test_gpu_lu.py), on the oracle's factors: equal IPIV passes; an IPIV that
differs at a step without a tie (a planted mistake) fails, traced in either
set of factors; the row tracking follows the interchange sequence; the
tolerance is the componentwise rounding bound and grows with the step and
the pivot growth."""
import numpy as np
import pytest

import lapack_amd as la
import oracle
from test_gpu_lu import assert_ipiv, final_row, tie_tolerance


def test_equal_ipiv_passes():
    a = la.matrix(2, 11, 40, 40)
    lu, p, _ = oracle.dgetrf(a)
    assert_ipiv(p, p, lu, lu)


@pytest.mark.parametrize("k", [0, 5, 20])
def test_planted_divergence_fails(k):
    n = 40
    a = la.matrix(2, 12 + k, n, n)
    lu, p, _ = oracle.dgetrf(a)
    bad = p.copy()
    bad[k] = n if p[k] != n else n - 1   # another row, no tie there
    with pytest.raises(AssertionError):
        assert_ipiv(bad, p, lu)
    with pytest.raises(AssertionError):
        assert_ipiv(bad, p, lu, lu)


def test_final_row_follows_swaps():
    # IPIV (1-based) [2, 4, 4, 4, 5]; 0-based interchanges k <-> [1, 3, 3, 3, 4]
    ipiv = np.array([2, 4, 4, 4, 5], dtype=np.int32)
    # row 3 at step 1: step 1 swaps 1 <-> 3 -> row 1; later steps leave row 1
    assert final_row(ipiv, 1, 3) == 1
    # row 1 at step 1: -> row 3; step 2 swaps 2 <-> 3 -> row 2; step 3 is 3 <-> 3
    assert final_row(ipiv, 1, 1) == 2
    # row 2 at step 1: untouched by step 1; step 2 moves it to row 3
    assert final_row(ipiv, 1, 2) == 3
    # brute force against the permuted row indices
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = 12
        ip = np.array([rng.integers(k, n) + 1 for k in range(n)], dtype=np.int32)
        k = int(rng.integers(0, n))
        rows = list(range(n))  # rows[i] = content label at position i
        for j in range(k, n):
            q = ip[j] - 1
            rows[j], rows[q] = rows[q], rows[j]
        for r in range(n):
            assert rows[final_row(ip, k, r)] == r


def test_tolerance_is_the_rounding_bound():
    n = 64
    a = la.matrix(2, 13, n, n)
    lu, _, _ = oracle.dgetrf(a)
    t0 = tie_tolerance(lu, 0, 2.0 ** -53)
    assert t0 == pytest.approx(10 * 2.0 ** -53 / (1 - 2.0 ** -53), rel=1e-12)  # step 0: exact entries, g = |u_00|
    t32 = tie_tolerance(lu, 32, 2.0 ** -53)
    assert t32 > t0
    assert tie_tolerance(lu, 32, 2.0 ** -24) > 1e8 * t32
