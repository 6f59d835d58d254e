// gemv.hip -- skinny FP64 products of the refinement loops, bound by HBM:
//   Y <- alpha*A*X + beta*Y   with A m x n, X n x r, r <= 16  (the residual
//                             R = B - A*X of SRC/dsgesv.f:328-332 and GMRES's
//                             products with A and with its basis V)
//   h <- alpha*V^T*w + beta*h with V m x k, k <= 32         (GMRES's
//                             projections and norms)
// Both read the big operand once, coalesced, and split the long dimension
// over workgroups; partial sums go to a workspace and a second kernel adds
// them in a fixed order, so results are deterministic run to run.
// beta == 0 does not read the output (BLAS semantics).
#include "la_common.h"

namespace la {
namespace {

constexpr int GV_ROWS = 256;   // rows per workgroup (one per thread)
constexpr int GV_CHUNK = 256;  // columns of X staged in LDS at a time

template <int R, bool ABS>
__global__ __launch_bounds__(256) void gemv_n_partial_kernel(int m, int n, int r, int cols_per_split,
                                                             const double* __restrict__ A, long lda,
                                                             const double* __restrict__ X, long ldx,
                                                             double* __restrict__ part) {
  __shared__ double sx[GV_CHUNK][R];
  const int i = blockIdx.x * GV_ROWS + threadIdx.x;
  const int s = blockIdx.y;
  const int j0 = s * cols_per_split;
  const int j1 = min(n, j0 + cols_per_split);
  double acc[R];
#pragma unroll
  for (int q = 0; q < R; ++q) acc[q] = 0.0;
  for (int jc = j0; jc < j1; jc += GV_CHUNK) {
    const int cn = min(GV_CHUNK, j1 - jc);
    __syncthreads();
    for (int e = threadIdx.x; e < GV_CHUNK * R; e += 256) {
      const int j = e % GV_CHUNK, q = e / GV_CHUNK;
      const double xv = (j < cn && q < r) ? X[(long)(jc + j) + (long)q * ldx] : 0.0;
      sx[j][q] = ABS ? fabs(xv) : xv;
    }
    __syncthreads();
    if (i < m) {
      const double* ap = A + i + (long)jc * lda;
      int j = 0;
      for (; j + 8 <= cn; j += 8) {
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = ABS ? fabs(ap[(long)(j + u) * lda]) : ap[(long)(j + u) * lda];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int q = 0; q < R; ++q) acc[q] = fma(a[u], sx[j + u][q], acc[q]);
      }
      for (; j < cn; ++j) {
        const double a = ABS ? fabs(ap[(long)j * lda]) : ap[(long)j * lda];
#pragma unroll
        for (int q = 0; q < R; ++q) acc[q] = fma(a, sx[j][q], acc[q]);
      }
    }
  }
  if (i < m) {
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (q < r) part[((long)s * r + q) * m + i] = acc[q];
  }
}

__global__ void gemv_n_reduce_kernel(int m, int r, int nsplit, double alpha, const double* __restrict__ part,
                                     double beta, double* __restrict__ Y, long ldy) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)m * r) return;
  const int i = (int)(e % m), q = (int)(e / m);
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[((long)k * r + q) * m + i];
  double* yp = Y + i + (long)q * ldy;
  *yp = (beta == 0.0) ? alpha * s : alpha * s + beta * *yp;
}

constexpr int GT_ROWS = 1024;  // rows per workgroup (four per thread)

template <int KC>
__global__ __launch_bounds__(256) void gemv_t_partial_kernel(int m, int k, const double* __restrict__ V, long ldv,
                                                             const double* __restrict__ w,
                                                             double* __restrict__ part) {
  __shared__ double red[4][KC];
  double acc[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) acc[c] = 0.0;
  const int base = blockIdx.x * GT_ROWS;
#pragma unroll
  for (int u = 0; u < GT_ROWS / 256; ++u) {
    const int i = base + u * 256 + threadIdx.x;
    if (i < m) {
      const double wi = w[i];
#pragma unroll
      for (int c = 0; c < KC; ++c)
        if (c < k) acc[c] = fma(V[i + (long)c * ldv], wi, acc[c]);
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    double v = acc[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < KC && threadIdx.x < k) {
    const int c = threadIdx.x;
    part[(long)blockIdx.x * KC + c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
  }
}

template <int KC>
__global__ void gemv_t_reduce_kernel(int k, int nblk, double alpha, const double* __restrict__ part, double beta,
                                     double* __restrict__ h) {
  const int c = threadIdx.x;
  if (c >= k) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[(long)b * KC + c];
  h[c] = (beta == 0.0) ? alpha * s : alpha * s + beta * h[c];
}

// y(j) <- alpha * sum_i op(A(i,j)) x(i) + beta*y(j) for every column j of an
// m x n matrix (op = identity or |.|, applied to x too): one wavefront per
// column, reading it contiguously; lanes add in a fixed order.
template <bool ABS>
__global__ __launch_bounds__(256) void gemv_cols_kernel(int m, int n, double alpha, const double* __restrict__ A,
                                                        long lda, const double* __restrict__ x, double beta,
                                                        double* __restrict__ y) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;
  const double* col = A + (long)j * lda;
  double s = 0.0;
  for (int i = lane; i < m; i += 64) {
    const double a = col[i], xv = x[i];
    s = fma(ABS ? fabs(a) : a, ABS ? fabs(xv) : xv, s);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) y[j] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[j];
}

int gemv_n_splits(int m, int n) {
  const int rb = (m + GV_ROWS - 1) / GV_ROWS;
  int s = std::max(1, 2048 / std::max(rb, 1));
  s = std::min(s, std::max(1, n / GV_CHUNK));
  return std::min(s, 64);
}

}  // namespace

size_t gemv_ws_doubles(int m, int n, int r) {
  const size_t a = (size_t)gemv_n_splits(m, n) * (size_t)std::max(r, 1) * (size_t)std::max(m, 1);
  const size_t b = (size_t)((m + GT_ROWS - 1) / GT_ROWS) * 32;
  return std::max(a, b);
}

void launch_dgemv_n(int m, int n, int r, double alpha, const double* A, long lda, const double* X, long ldx,
                    double beta, double* Y, long ldy, double* ws, hipStream_t s) {
  if (m <= 0 || r <= 0) return;
  const int nsplit = n > 0 ? gemv_n_splits(m, n) : 1;
  const int cps = n > 0 ? (n + nsplit - 1) / nsplit : 0;
  if (n > 0) {
    dim3 g((m + GV_ROWS - 1) / GV_ROWS, nsplit);
    if (r <= 1) gemv_n_partial_kernel<1, false><<<g, 256, 0, s>>>(m, n, r, cps, A, lda, X, ldx, ws);
    else if (r <= 2) gemv_n_partial_kernel<2, false><<<g, 256, 0, s>>>(m, n, r, cps, A, lda, X, ldx, ws);
    else if (r <= 4) gemv_n_partial_kernel<4, false><<<g, 256, 0, s>>>(m, n, r, cps, A, lda, X, ldx, ws);
    else if (r <= 8) gemv_n_partial_kernel<8, false><<<g, 256, 0, s>>>(m, n, r, cps, A, lda, X, ldx, ws);
    else gemv_n_partial_kernel<16, false><<<g, 256, 0, s>>>(m, n, r, cps, A, lda, X, ldx, ws);
  } else {
    launch_fill_i32(reinterpret_cast<int*>(ws), (size_t)2 * m * r, 0, s);
  }
  const long tot = (long)m * r;
  gemv_n_reduce_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(m, r, n > 0 ? nsplit : 1, alpha, ws, beta, Y,
                                                                     ldy);
}

// y <- alpha*A^T*x + beta*y (A m x n, y has n entries); abs: |A|^T |x|.
void launch_dgemv_cols(int m, int n, double alpha, const double* A, long lda, const double* x, double beta,
                       double* y, bool abs, hipStream_t s) {
  if (n <= 0) return;
  const int g = (n + 3) / 4;
  if (abs) gemv_cols_kernel<true><<<g, 256, 0, s>>>(m, n, alpha, A, lda, x, beta, y);
  else gemv_cols_kernel<false><<<g, 256, 0, s>>>(m, n, alpha, A, lda, x, beta, y);
}

// Y <- alpha*|A|*|X| + beta*Y for one column X (the |A||x| term of DGERFS).
void launch_dgemv_n_abs(int m, int n, double alpha, const double* A, long lda, const double* X, double beta,
                        double* Y, double* ws, hipStream_t s) {
  if (m <= 0) return;
  const int nsplit = n > 0 ? gemv_n_splits(m, n) : 1;
  const int cps = n > 0 ? (n + nsplit - 1) / nsplit : 0;
  if (n > 0) {
    dim3 g((m + GV_ROWS - 1) / GV_ROWS, nsplit);
    gemv_n_partial_kernel<1, true><<<g, 256, 0, s>>>(m, n, 1, cps, A, lda, X, m, ws);
  } else {
    launch_fill_i32(reinterpret_cast<int*>(ws), (size_t)2 * m, 0, s);
  }
  gemv_n_reduce_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(m, 1, n > 0 ? nsplit : 1, alpha, ws, beta, Y,
                                                                   m);
}

void launch_dgemv_t(int m, int k, double alpha, const double* V, long ldv, const double* w, double beta, double* h,
                    double* ws, hipStream_t s) {
  if (k <= 0) return;
  const int nblk = std::max(1, (m + GT_ROWS - 1) / GT_ROWS);
  if (k <= 1) {
    gemv_t_partial_kernel<1><<<nblk, 256, 0, s>>>(m, k, V, ldv, w, ws);
    gemv_t_reduce_kernel<1><<<1, 64, 0, s>>>(k, nblk, alpha, ws, beta, h);
  } else if (k <= 8) {
    gemv_t_partial_kernel<8><<<nblk, 256, 0, s>>>(m, k, V, ldv, w, ws);
    gemv_t_reduce_kernel<8><<<1, 64, 0, s>>>(k, nblk, alpha, ws, beta, h);
  } else if (k <= 16) {
    gemv_t_partial_kernel<16><<<nblk, 256, 0, s>>>(m, k, V, ldv, w, ws);
    gemv_t_reduce_kernel<16><<<1, 64, 0, s>>>(k, nblk, alpha, ws, beta, h);
  } else {
    gemv_t_partial_kernel<32><<<nblk, 256, 0, s>>>(m, k, V, ldv, w, ws);
    gemv_t_reduce_kernel<32><<<1, 64, 0, s>>>(k, nblk, alpha, ws, beta, h);
  }
}

}  // namespace la
