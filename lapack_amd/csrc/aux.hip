// aux.hip -- data-movement kernels of the LU path: transposes into and out of
// the factorization workspace, row interchanges (DLASWP, SRC/dlaswp.f:112-189),
// interchange-sequence -> permutation conversion for DGETRS, triangular
// block inverses used by the blocked triangular solves, copies and fills.
#include "la_common.h"

namespace la {
namespace {

// ---- transpose ------------------------------------------------------------
// out(j,i) = cast(in(i,j)), 32x32 tiles through LDS (+1 padding against bank
// conflicts), reads and writes both coalesced.
template <class Ti, class To, bool CHECK>
__global__ __launch_bounds__(256) void transpose_kernel(int rows, int cols, const Ti* __restrict__ in,
                                                        long ld_in, To* __restrict__ out, long ld_out,
                                                        int* overflow) {
  __shared__ Ti tile[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
  for (int r = 0; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + ty + r;
    if (i < rows && j < cols) tile[ty + r][tx] = in[(long)i + (long)j * ld_in];
  }
  __syncthreads();
  bool ovf = false;
#pragma unroll
  for (int r = 0; r < 32; r += 8) {
    const int j = j0 + tx, i = i0 + ty + r;
    if (i < rows && j < cols) {
      const Ti v = tile[tx][ty + r];
      if (CHECK) {
        // DLAG2S rejects |a| > SLAMCH('O') (SRC/dlag2s.f:135-139).
        if (v < -(Ti)3.4028234663852886e38 || v > (Ti)3.4028234663852886e38) ovf = true;
      }
      out[(long)j + (long)i * ld_out] = (To)v;
    }
  }
  if (CHECK && ovf) atomicOr(overflow, 1);
}

template <class Ti, class To, bool CHECK>
void launch_transpose_t(int rows, int cols, const Ti* in, long ld_in, To* out, long ld_out, int* ovf,
                        hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  dim3 grid((rows + 31) / 32, (cols + 31) / 32);
  transpose_kernel<Ti, To, CHECK><<<grid, 256, 0, s>>>(rows, cols, in, ld_in, out, ld_out, ovf);
}

// ---- row interchanges in the transposed workspace ---------------------------
// T = A^T: interchanging rows r and p of A is interchanging columns r and p of
// T.  One thread per T row c in [r0, r1); the interchanges are applied in the
// order of SRC/dlaswp.f:156-166 (k = k1..k2-1), so a row hit twice ends where
// the reference puts it.  Consecutive threads touch consecutive addresses.
template <class T>
__global__ __launch_bounds__(256) void laswp_cols_kernel(T* __restrict__ Tm, long ldt, int r0, int r1,
                                                         int k1, int k2, const int* __restrict__ ipiv) {
  __shared__ int sp[512];
  const int nk = k2 - k1;
  for (int i = threadIdx.x; i < nk; i += 256) sp[i] = ipiv[k1 + i] - 1;
  __syncthreads();
  const long c = (long)r0 + (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= r1) return;
  for (int i = 0; i < nk; ++i) {
    const int r = k1 + i, p = sp[i];
    if (p != r) {
      T* a = Tm + c + (long)r * ldt;
      T* b = Tm + c + (long)p * ldt;
      const T t = *a;
      *a = *b;
      *b = t;
    }
  }
}

// ---- DLASWP on a column-major matrix (row interchanges) ----------------------
// Used on the right-hand sides of DGETRS (SRC/dgetrs.f:187,217) and by the
// dlaswp_ entry point.  The interchange sequence is first turned into a
// permutation (perm[r] = the row whose content ends at row r), walked once
// in the reference's order (SRC/dlaswp.f:138-150: K1..K2 for INCX > 0, K2..K1
// for INCX < 0) over a position map held in LDS; the rows are then gathered
// through it in parallel.
constexpr int PERM_LDS_MAX = 32768;
constexpr int PERM_CHUNK = 1024;

template <bool LDSMAP>
__global__ __launch_bounds__(256) void swaps_to_perm_kernel(int nrows, int k1, int k2,
                                                            const int* __restrict__ ipiv, int incx,
                                                            int* __restrict__ perm) {
  extern __shared__ int smem[];
  int* sp = smem;                                  // PERM_CHUNK staged pivots
  int* cur = LDSMAP ? smem + PERM_CHUNK : perm;    // position map
  for (int r = threadIdx.x; r < nrows; r += 256) cur[r] = r;
  __syncthreads();
  const int nk = k2 - k1;
  for (int c0 = 0; c0 < nk; c0 += PERM_CHUNK) {
    const int cn = min(PERM_CHUNK, nk - c0);
    // chunk c0 in processing order: forward k1+c0.., reverse k2-1-c0 downwards
    for (int t = threadIdx.x; t < cn; t += 256) {
      const int i = incx > 0 ? k1 + c0 + t : k2 - 1 - c0 - t;
      sp[t] = ipiv[i] - 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int t = 0; t < cn; ++t) {
        const int i = incx > 0 ? k1 + c0 + t : k2 - 1 - c0 - t;
        const int p = sp[t];
        if (p != i && p >= 0 && p < nrows) {  // out-of-range entries are ignored
          const int a = cur[i];
          cur[i] = cur[p];
          cur[p] = a;
        }
      }
    }
    __syncthreads();
  }
  if (LDSMAP)
    for (int r = threadIdx.x; r < nrows; r += 256) perm[r] = cur[r];
}

// dst(r,j) = src(perm[r], j) for r < nrows, j < ncols.
template <class T>
__global__ __launch_bounds__(256) void gather_rows_kernel(int nrows, int ncols, const int* __restrict__ perm,
                                                          const T* __restrict__ src, long lds,
                                                          T* __restrict__ dst, long ldd) {
  const long tot = (long)nrows * ncols;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const int r = (int)(e % nrows), j = (int)(e / nrows);
    dst[(long)r + (long)j * ldd] = src[(long)perm[r] + (long)j * lds];
  }
}

// ---- triangular block inverses -------------------------------------------
// W = inv(tri(U)) for nb x nb blocks, nb <= 256; blockIdx.x = column c of W,
// blockIdx.y = which diagonal block (batched along the diagonal with stride
// nb: block d starts at U + d*nb*(ldu+1), W + d*nb*ldw).  One thread per row.
// UPPER: solve tri(U) x = e_c by backward substitution (column-oriented, the
// loop order of BLAS/SRC/dtrsm.f:277-288 on a single right-hand side);
// lower: forward substitution (dtrsm.f:289-300).  UNIT: unit diagonal.
template <class T, bool UPPER, bool UNIT>
__global__ __launch_bounds__(256) void trtri_kernel(int n, int nb, const T* __restrict__ U, long ldu,
                                                    T* __restrict__ W, long ldw) {
  __shared__ T xs;
  const int c = blockIdx.x;
  const int d = blockIdx.y;
  const int b0 = d * nb;
  const int bs = min(nb, n - b0);
  if (c >= bs) return;
  const T* Ub = U + (long)b0 * (ldu + 1);
  T* Wb = W + (long)d * nb * ldw;  // W block d: nb x nb at column offset d*nb
  const int s = threadIdx.x;
  T x = (s == c) ? T(1) : T(0);
  if (UPPER) {
    for (int t = c; t >= 0; --t) {
      if (s == t) {
        if (!UNIT) x = x / Ub[(long)t + (long)t * ldu];
        xs = x;
      }
      __syncthreads();
      const T xt = xs;
      if (s < t) x = x - xt * Ub[(long)s + (long)t * ldu];
      __syncthreads();
    }
  } else {
    for (int t = c; t < bs; ++t) {
      if (s == t) {
        if (!UNIT) x = x / Ub[(long)t + (long)t * ldu];
        xs = x;
      }
      __syncthreads();
      const T xt = xs;
      if (s > t && s < bs) x = x - xt * Ub[(long)s + (long)t * ldu];
      __syncthreads();
    }
  }
  if (s < bs) Wb[(long)s + (long)c * ldw] = x;
}

// ---- copies and fills -----------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void copy_kernel(int rows, int cols, const T* __restrict__ src, long lds,
                                                   T* __restrict__ dst, long ldd) {
  const long tot = (long)rows * cols;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const long i = e % rows, j = e / rows;
    dst[i + j * ldd] = src[i + j * lds];
  }
}
__global__ void fold_err_kernel(const int* __restrict__ err, int* __restrict__ info) {
  if (threadIdx.x == 0 && *err) *info = LA_ERR_SYNC;
}

template <class T>
__global__ __launch_bounds__(256) void fill_kernel(T* __restrict__ p, size_t n, T v) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) p[e] = v;
}

inline int grid_for(long work) {
  long g = (work + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

void launch_transpose_d(int rows, int cols, const double* in, long ld_in, double* out, long ld_out,
                        hipStream_t s) {
  launch_transpose_t<double, double, false>(rows, cols, in, ld_in, out, ld_out, nullptr, s);
}
void launch_transpose_s(int rows, int cols, const float* in, long ld_in, float* out, long ld_out,
                        hipStream_t s) {
  launch_transpose_t<float, float, false>(rows, cols, in, ld_in, out, ld_out, nullptr, s);
}
void launch_transpose_d2s(int rows, int cols, const double* in, long ld_in, float* out, long ld_out,
                          int* overflow, hipStream_t s) {
  launch_transpose_t<double, float, true>(rows, cols, in, ld_in, out, ld_out, overflow, s);
}

void launch_laswp_cols_d(double* T, long ldt, int r0, int r1, int k1, int k2, const int* ipiv,
                         hipStream_t s) {
  for (int kb = k1; kb < k2; kb += 512) {
    const int ke = kb + 512 < k2 ? kb + 512 : k2;
    if (r1 > r0) laswp_cols_kernel<double><<<(r1 - r0 + 255) / 256, 256, 0, s>>>(T, ldt, r0, r1, kb, ke, ipiv);
  }
}
void launch_laswp_cols_s(float* T, long ldt, int r0, int r1, int k1, int k2, const int* ipiv,
                         hipStream_t s) {
  for (int kb = k1; kb < k2; kb += 512) {
    const int ke = kb + 512 < k2 ? kb + 512 : k2;
    if (r1 > r0) laswp_cols_kernel<float><<<(r1 - r0 + 255) / 256, 256, 0, s>>>(T, ldt, r0, r1, kb, ke, ipiv);
  }
}

// perm (nrows ints) <- permutation of the interchange sequence; see above.
void launch_swaps_to_perm(int nrows, int k1, int k2, const int* ipiv, int incx, int* perm, hipStream_t s) {
  static int lds_max = -1;
  if (lds_max < 0) {
    int dev = 0, v = 65536;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
    lds_max = v;
  }
  const size_t need = sizeof(int) * (PERM_CHUNK + (size_t)nrows);
  if (nrows <= PERM_LDS_MAX && need <= (size_t)lds_max) {
    const size_t sh = sizeof(int) * (PERM_CHUNK + (size_t)nrows);
    swaps_to_perm_kernel<true><<<1, 256, sh, s>>>(nrows, k1, k2, ipiv, incx, perm);
  } else {
    swaps_to_perm_kernel<false><<<1, 256, sizeof(int) * PERM_CHUNK, s>>>(nrows, k1, k2, ipiv, incx, perm);
  }
}
void launch_gather_rows_d(int nrows, int ncols, const int* perm, const double* src, long lds, double* dst,
                          long ldd, hipStream_t s) {
  if (nrows > 0 && ncols > 0)
    gather_rows_kernel<double><<<grid_for((long)nrows * ncols), 256, 0, s>>>(nrows, ncols, perm, src, lds, dst, ldd);
}
void launch_gather_rows_s(int nrows, int ncols, const int* perm, const float* src, long lds, float* dst,
                          long ldd, hipStream_t s) {
  if (nrows > 0 && ncols > 0)
    gather_rows_kernel<float><<<grid_for((long)nrows * ncols), 256, 0, s>>>(nrows, ncols, perm, src, lds, dst, ldd);
}

template <class T>
static void launch_trtri_t(bool upper, bool unit, int n, int nb, const T* U, long ldu, T* W, long ldw,
                           hipStream_t s) {
  if (n <= 0) return;
  const int nblk = (n + nb - 1) / nb;
  dim3 grid(nb, nblk);
  if (upper && unit) trtri_kernel<T, true, true><<<grid, 256, 0, s>>>(n, nb, U, ldu, W, ldw);
  else if (upper) trtri_kernel<T, true, false><<<grid, 256, 0, s>>>(n, nb, U, ldu, W, ldw);
  else if (unit) trtri_kernel<T, false, true><<<grid, 256, 0, s>>>(n, nb, U, ldu, W, ldw);
  else trtri_kernel<T, false, false><<<grid, 256, 0, s>>>(n, nb, U, ldu, W, ldw);
}
void launch_trtri_blocks_d(bool upper, bool unit, int n, int nb, const double* U, long ldu, double* W,
                           long ldw, hipStream_t s) {
  launch_trtri_t<double>(upper, unit, n, nb, U, ldu, W, ldw, s);
}
void launch_trtri_blocks_s(bool upper, bool unit, int n, int nb, const float* U, long ldu, float* W,
                           long ldw, hipStream_t s) {
  launch_trtri_t<float>(upper, unit, n, nb, U, ldu, W, ldw, s);
}
void launch_trtri_uu_d(int nb, const double* U, long ldu, double* W, long ldw, hipStream_t s) {
  launch_trtri_t<double>(true, true, nb, nb, U, ldu, W, ldw, s);
}
void launch_trtri_uu_s(int nb, const float* U, long ldu, float* W, long ldw, hipStream_t s) {
  launch_trtri_t<float>(true, true, nb, nb, U, ldu, W, ldw, s);
}
void launch_trtri_un_d(int nb, const double* U, long ldu, double* W, long ldw, hipStream_t s) {
  launch_trtri_t<double>(true, false, nb, nb, U, ldu, W, ldw, s);
}
void launch_trtri_un_s(int nb, const float* U, long ldu, float* W, long ldw, hipStream_t s) {
  launch_trtri_t<float>(true, false, nb, nb, U, ldu, W, ldw, s);
}

void launch_copy_d(int rows, int cols, const double* src, long lds, double* dst, long ldd, hipStream_t s) {
  if (rows > 0 && cols > 0) copy_kernel<double><<<grid_for((long)rows * cols), 256, 0, s>>>(rows, cols, src, lds, dst, ldd);
}
void launch_copy_s(int rows, int cols, const float* src, long lds, float* dst, long ldd, hipStream_t s) {
  if (rows > 0 && cols > 0) copy_kernel<float><<<grid_for((long)rows * cols), 256, 0, s>>>(rows, cols, src, lds, dst, ldd);
}
void launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s) {
  if (n) fill_kernel<unsigned><<<grid_for((long)n), 256, 0, s>>>(p, n, v);
}
void launch_fold_err(const int* err, int* info, hipStream_t s) { fold_err_kernel<<<1, 64, 0, s>>>(err, info); }
void launch_fill_i32(int* p, size_t n, int v, hipStream_t s) {
  if (n) fill_kernel<int><<<grid_for((long)n), 256, 0, s>>>(p, n, v);
}

}  // namespace la
