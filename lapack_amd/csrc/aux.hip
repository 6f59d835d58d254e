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
// T.  The panel's interchange sequence ipiv[k1..k2) (SRC/dlaswp.f:156-166 order)
// is first composed into (dst, src) pairs over the positions it touches --
// one workgroup walks the sequence once over a position map in LDS -- and the
// columns are then moved in a single pass: each workgroup owns 16 T rows (one
// 128-byte line per T column), stages every source value in LDS, barriers,
// and writes every destination.  Two row ranges (left and right of the panel)
// are served by one launch.
constexpr int SWP_ROWS = 16;      // T rows per workgroup
constexpr int SWP_MAXPAIRS = 512; // 2 * the largest panel width

// pairs[2q] = dst, pairs[2q+1] = src (0-based positions); *npairs = count.
// Positions live in [k1, npos); map is LDS (or global when it does not fit).
template <bool LDSMAP>
__global__ __launch_bounds__(256) void swaps_to_pairs_kernel(int k1, int k2, int npos, const int* __restrict__ ipiv,
                                                             int* __restrict__ pairs, int* __restrict__ npairs,
                                                             int* __restrict__ gmap) {
  extern __shared__ int smem[];
  __shared__ int s_cnt;
  int* cur = LDSMAP ? smem : gmap;
  const int span = npos - k1;
  for (int x = threadIdx.x; x < span; x += 256) cur[x] = k1 + x;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = k1; i < k2; ++i) {
      const int p = ipiv[i] - 1;
      if (p != i && p >= k1 && p < npos) {
        const int t = cur[i - k1];
        cur[i - k1] = cur[p - k1];
        cur[p - k1] = t;
      }
    }
  }
  __syncthreads();
  for (int x = threadIdx.x; x < span; x += 256) {
    const int src = cur[x];
    if (src != k1 + x) {
      const int q = atomicAdd(&s_cnt, 1);
      if (q < SWP_MAXPAIRS) {
        pairs[2 * q] = k1 + x;
        pairs[2 * q + 1] = src;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *npairs = min(s_cnt, SWP_MAXPAIRS);
}

template <class T>
__global__ __launch_bounds__(256) void apply_pairs_cols_kernel(T* __restrict__ Tm, long ldt, int r0, int r1, int r2,
                                                               int r3, const int* __restrict__ pairs,
                                                               const int* __restrict__ npairs) {
  __shared__ T st[SWP_MAXPAIRS * SWP_ROWS];
  const int np = *npairs;
  // this block's 16 rows: first from [r0, r1), then from [r2, r3)
  const int n1 = (r1 - r0 + SWP_ROWS - 1) / SWP_ROWS;
  const int blk = blockIdx.x;
  const long c0 = blk < n1 ? (long)r0 + (long)blk * SWP_ROWS : (long)r2 + (long)(blk - n1) * SWP_ROWS;
  const long cend = blk < n1 ? r1 : r3;
  const int cl = threadIdx.x & (SWP_ROWS - 1);
  const int qg = threadIdx.x / SWP_ROWS;  // 16 pair lanes
  const long c = c0 + cl;
  const bool ok = c < cend;
  for (int q = qg; q < np; q += 256 / SWP_ROWS)
    if (ok) st[q * SWP_ROWS + cl] = Tm[c + (long)pairs[2 * q + 1] * ldt];
  __syncthreads();
  for (int q = qg; q < np; q += 256 / SWP_ROWS)
    if (ok) Tm[c + (long)pairs[2 * q] * ldt] = st[q * SWP_ROWS + cl];
}

// Row interchanges of a column-major matrix through the same pairs:
// B(dst, j) <- B_old(src, j); one workgroup per column block of 16.
template <class T>
__global__ __launch_bounds__(256) void apply_pairs_rows_kernel(T* __restrict__ B, long ldb, int ncols,
                                                               const int* __restrict__ pairs,
                                                               const int* __restrict__ npairs) {
  __shared__ T st[SWP_MAXPAIRS * SWP_ROWS];
  const int np = *npairs;
  const int j = blockIdx.x * SWP_ROWS + (threadIdx.x & (SWP_ROWS - 1));
  const int qg = threadIdx.x / SWP_ROWS;
  const bool ok = j < ncols;
  for (int q = qg; q < np; q += 256 / SWP_ROWS)
    if (ok) st[q * SWP_ROWS + (threadIdx.x & (SWP_ROWS - 1))] = B[(long)pairs[2 * q + 1] + (long)j * ldb];
  __syncthreads();
  for (int q = qg; q < np; q += 256 / SWP_ROWS)
    if (ok) B[(long)pairs[2 * q] + (long)j * ldb] = st[q * SWP_ROWS + (threadIdx.x & (SWP_ROWS - 1))];
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

int lds_limit() {
  static int lds_max = -1;
  if (lds_max < 0) {
    int dev = 0, v = 65536;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
    lds_max = v;
  }
  return lds_max;
}

int swp_max_pairs() { return SWP_MAXPAIRS; }

// (dst, src) pairs of the interchange sequence ipiv[k1..k2) over positions
// [k1, npos); gmap: npos ints of scratch used when the map does not fit LDS.
void launch_swaps_to_pairs(int k1, int k2, int npos, const int* ipiv, int* pairs, int* npairs, int* gmap,
                           hipStream_t s) {
  const size_t need = sizeof(int) * (size_t)(npos - k1) + 64;
  if (need <= (size_t)lds_limit())
    swaps_to_pairs_kernel<true><<<1, 256, sizeof(int) * (size_t)(npos - k1), s>>>(k1, k2, npos, ipiv, pairs, npairs, gmap);
  else
    swaps_to_pairs_kernel<false><<<1, 256, 0, s>>>(k1, k2, npos, ipiv, pairs, npairs, gmap);
}
template <class T>
static void apply_cols_t(T* Tm, long ldt, int r0, int r1, int r2, int r3, const int* pairs, const int* npairs,
                         hipStream_t s) {
  if (r1 < r0) r1 = r0;
  if (r3 < r2) r3 = r2;
  const int nb = (r1 - r0 + SWP_ROWS - 1) / SWP_ROWS + (r3 - r2 + SWP_ROWS - 1) / SWP_ROWS;
  if (nb > 0) apply_pairs_cols_kernel<T><<<nb, 256, 0, s>>>(Tm, ldt, r0, r1, r2, r3, pairs, npairs);
}
void launch_apply_pairs_cols_d(double* T, long ldt, int r0, int r1, int r2, int r3, const int* pairs,
                               const int* npairs, hipStream_t s) {
  apply_cols_t<double>(T, ldt, r0, r1, r2, r3, pairs, npairs, s);
}
void launch_apply_pairs_cols_s(float* T, long ldt, int r0, int r1, int r2, int r3, const int* pairs,
                               const int* npairs, hipStream_t s) {
  apply_cols_t<float>(T, ldt, r0, r1, r2, r3, pairs, npairs, s);
}
void launch_apply_pairs_rows_d(double* B, long ldb, int ncols, const int* pairs, const int* npairs, hipStream_t s) {
  if (ncols > 0) apply_pairs_rows_kernel<double><<<(ncols + SWP_ROWS - 1) / SWP_ROWS, 256, 0, s>>>(B, ldb, ncols, pairs, npairs);
}
void launch_apply_pairs_rows_s(float* B, long ldb, int ncols, const int* pairs, const int* npairs, hipStream_t s) {
  if (ncols > 0) apply_pairs_rows_kernel<float><<<(ncols + SWP_ROWS - 1) / SWP_ROWS, 256, 0, s>>>(B, ldb, ncols, pairs, npairs);
}

// perm (nrows ints) <- permutation of the interchange sequence; see above.
void launch_swaps_to_perm(int nrows, int k1, int k2, const int* ipiv, int incx, int* perm, hipStream_t s) {
  const size_t need = sizeof(int) * (PERM_CHUNK + (size_t)nrows);
  if (nrows <= PERM_LDS_MAX && need <= (size_t)lds_limit()) {
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
