// adjoint.hip -- the rank-nrhs outer-product kernel of the solve's adjoint.
//
// For X = A^-1 B with loss gradient Xbar, the reverse pass is
//   Bbar = A^-T Xbar          (DGETRS('T') on the saved LU, SRC/dgetrs.f:199-217)
//   Abar = -Bbar X^T          (a DGEMM('N','T') with K = NRHS, BLAS/SRC/dgemm.f)
// The second product has tiny K (NRHS) and an n x n result: it is bound by
// the 8 n^2 bytes written to HBM, not by arithmetic, so it is a plain
// vector-ALU kernel: a 64 x 64 tile of C per 256-thread workgroup, the two
// 64 x K strips staged in LDS, each lane writing 16 columns of one row so
// every store of a wavefront is 512 contiguous bytes of a column.
#include "la_common.h"

namespace la {
namespace {

constexpr int RK_TILE = 64;
constexpr int RK_KMAX = 32;

// C(i,j) <- alpha * sum_p A(i,p) B(j,p) + beta * C(i,j); beta == 0 does not
// read C (BLAS semantics: NaN in C is not propagated).
__global__ __launch_bounds__(256) void rankk_nt_kernel(int m, int n, int k, double alpha, const double* __restrict__ A,
                                                       long lda, const double* __restrict__ B, long ldb, double beta,
                                                       double* __restrict__ C, long ldc) {
  __shared__ double sa[RK_KMAX][RK_TILE];
  __shared__ double sb[RK_KMAX][RK_TILE + 1];
  const int i0 = blockIdx.x * RK_TILE, j0 = blockIdx.y * RK_TILE;
  const int t = threadIdx.x;
  for (int e = t; e < k * RK_TILE; e += 256) {
    const int p = e / RK_TILE, r = e % RK_TILE;
    sa[p][r] = (i0 + r < m) ? A[(i0 + r) + (long)p * lda] : 0.0;
    sb[p][r] = (j0 + r < n) ? B[(j0 + r) + (long)p * ldb] : 0.0;
  }
  __syncthreads();
  const int r = t % RK_TILE, cg = t / RK_TILE;  // row r, columns cg*16 .. cg*16+15
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  for (int p = 0; p < k; ++p) {
    const double a = sa[p][r];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = fma(a, sb[p][cg * 16 + q], acc[q]);
  }
  const int i = i0 + r;
  if (i >= m) return;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int j = j0 + cg * 16 + q;
    if (j < n) {
      double* cp = C + i + (long)j * ldc;
      *cp = (beta == 0.0) ? alpha * acc[q] : alpha * acc[q] + beta * *cp;
    }
  }
}

}  // namespace

int rankk_max() { return RK_KMAX; }

void launch_rankk_nt_d(int m, int n, int k, double alpha, const double* A, long lda, const double* B, long ldb,
                       double beta, double* C, long ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  dim3 g((m + RK_TILE - 1) / RK_TILE, (n + RK_TILE - 1) / RK_TILE);
  rankk_nt_kernel<<<g, 256, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

}  // namespace la
