// gemm.hip -- C <- alpha*op(A)*op(B) + beta*C on the gfx950 matrix cores.
//
// This is the arithmetic of BLAS DGEMM/SGEMM (BLAS/SRC/dgemm.f:213-408) as
// used by the LU path: the trailing update DGEMM('N','N',...,-1,A21,U12,1,A22)
// of SRC/dgetrf.f:212-216 and SRC/dgetrf2.f:245-246, the residual
// R = B - A*X of SRC/dsgesv.f:328-332 and the rank-nrhs adjoint update.
// It is not a loop-order restatement: products are accumulated by
// v_mfma_f64_16x16x4_f64 (v_mfma_f32_16x16x4_f32 for FP32), so results agree
// with the reference to rounding, which the tests bound with the BLAS test
// suite's componentwise ratio (BLAS/TESTING/dblat3.f:2940-2960).
//
// Tiling: a 256-thread workgroup (4 waves, 2x2) owns a 128x128 tile of C;
// each wave owns 64x64 = 4x4 MFMA blocks held in 64 accumulator registers
// per lane.  K advances 16 at a time through double-buffered LDS:
//   As[k][i]  (k-major, i contiguous)   -- the MFMA A operand reads 16
//                                           consecutive i per 16 lanes
//   Bs[j][k]  (row stride BK+1)          -- conflict-free B operand reads
// Global->LDS loads are coalesced for either transpose of A and B.
// Workgroups are mapped XCD-aware: blockIdx % 8 selects the XCD (round-robin
// dispatch), and each XCD receives a contiguous, M-grouped range of tiles so
// the A row panels and B column panels it touches stay in its own L2.
#include "la_common.h"

namespace la {
namespace {

constexpr int BM = 128, BN = 128, BK = 16, NT = 256;
constexpr int GROUP_M = 8;
constexpr int NXCD = 8;

template <class T> struct Mfma;
template <> struct Mfma<double> {
  typedef la_d4 acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // Accumulator register r of lane l holds D[l/16 + 4r][l%16] (measured on
  // gfx950; the f32 form below differs).
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
  static constexpr int APAD = 8;  // (BM+APAD)*8 bytes = 64 mod 128
};
template <> struct Mfma<float> {
  typedef la_f4 acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // Accumulator register r of lane l holds D[4*(l/16) + r][l%16].
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
  static constexpr int APAD = 16;
};

// UPD: the trailing-update form C <- C - A*B (alpha = -1, beta = 1), a
// separate symbol so profiles can single out the DGETRF trailing update.
template <class T, bool TA, bool TB, bool UPD>
__global__ __launch_bounds__(NT, 2) void gemm_kernel(int M, int N, int K, T alpha,
                                                     const T* __restrict__ A, long lda,
                                                     const T* __restrict__ B, long ldb, T beta,
                                                     T* __restrict__ C, long ldc, int tm, int tn,
                                                     int chunk, int nxcd) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int AST = BM + Mfma<T>::APAD;  // As row stride
  constexpr int BST = BK + 1;              // Bs row stride
  __shared__ T As[2][BK * AST];
  __shared__ T Bs[2][BN * BST];

  // XCD-aware tile selection; with nxcd == 7 XCD 0 is left idle.
  const int bid = blockIdx.x;
  int xcd = bid % NXCD;
  if (nxcd < NXCD) {
    if (xcd < NXCD - nxcd) return;
    xcd -= NXCD - nxcd;
  }
  const int lin = xcd * chunk + bid / NXCD;
  if (lin >= tm * tn) return;
  const int width = GROUP_M * tn;
  const int group = lin / width;
  const int first_m = group * GROUP_M;
  const int gsize = min(tm - first_m, GROUP_M);
  const int tile_m = first_m + (lin % width) % gsize;
  const int tile_n = (lin % width) / gsize;
  const int row0 = tile_m * BM;
  const int col0 = tile_n * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1;   // wave row in the 2x2 wave grid
  const int wn = wave >> 1;  // wave column

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  T ra[8], rb[8];
  const int nkt = (K + BK - 1) / BK;

  auto load_tile = [&](int kt) {
    const int k0 = kt * BK;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int i, k;
      if (!TA) { i = tid & 127; k = (tid >> 7) + 2 * q; }
      else     { k = tid & 15;  i = (tid >> 4) + 16 * q; }
      const int gi = row0 + i, gk = k0 + k;
      T v = T(0);
      if (gi < M && gk < K) v = TA ? A[(long)gk + (long)gi * lda] : A[(long)gi + (long)gk * lda];
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int j, k;
      if (!TB) { k = tid & 15;  j = (tid >> 4) + 16 * q; }
      else     { j = tid & 127; k = (tid >> 7) + 2 * q; }
      const int gj = col0 + j, gk = k0 + k;
      T v = T(0);
      if (gj < N && gk < K) v = TB ? B[(long)gj + (long)gk * ldb] : B[(long)gk + (long)gj * ldb];
      rb[q] = v;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int i, k;
      if (!TA) { i = tid & 127; k = (tid >> 7) + 2 * q; }
      else     { k = tid & 15;  i = (tid >> 4) + 16 * q; }
      As[buf][k * AST + i] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int j, k;
      if (!TB) { k = tid & 15;  j = (tid >> 4) + 16 * q; }
      else     { j = tid & 127; k = (tid >> 7) + 2 * q; }
      Bs[buf][j * BST + k] = rb[q];
    }
  };

  if (nkt > 0) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();

  const int li = lane & 15;  // row (A) / column (B) inside a 16-block
  const int lk = lane >> 4;  // k inside a 4-step
  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nkt) load_tile(kt + 1);
    const T* as = As[cur];
    const T* bs = Bs[cur];
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      T af[4], bf[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        af[b] = as[(kk * 4 + lk) * AST + wm * 64 + b * 16 + li];
        bf[b] = bs[(wn * 64 + b * 16 + li) * BST + kk * 4 + lk];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < nkt) store_tile(cur ^ 1);
    __syncthreads();
  }

  // Epilogue: each lane holds four elements of column lane%16 of every
  // 16x16 block, at the rows given by Mfma<T>::row.
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gc = col0 + wn * 64 + j * 16 + li;
      if (gc >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wm * 64 + i * 16 + Mfma<T>::row(lk, r);
        if (gr < M) {
          T* cp = C + (long)gr + (long)gc * ldc;
          const T v = acc[i][j][r];
          if (UPD) *cp = *cp - v;
          else *cp = (beta == T(0)) ? alpha * v : alpha * v + beta * *cp;
        }
      }
    }
}

template <class T>
void launch_gemm_t(char ta, char tb, int m, int n, int k, T alpha, const T* A, long lda, const T* B,
                   long ldb, T beta, T* C, long ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  // BLAS quick return (BLAS/SRC/dgemm.f:298-299).
  if ((alpha == T(0) || k <= 0) && beta == T(1)) return;
  const int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
  const int tiles = tm * tn;
  const int chunk = (tiles + NXCD - 1) / NXCD;
  const dim3 grid(chunk * NXCD);
  const bool TA = (ta == 'T' || ta == 't' || ta == 'C' || ta == 'c');
  const bool TB = (tb == 'T' || tb == 't' || tb == 'C' || tb == 'c');
  if (k <= 0 || alpha == T(0)) k = 0;  // C <- beta*C (dgemm.f:303-318)
  if (!TA && !TB)
    gemm_kernel<T, false, false, false><<<grid, NT, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, chunk, NXCD);
  else if (!TA && TB)
    gemm_kernel<T, false, true, false><<<grid, NT, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, chunk, NXCD);
  else if (TA && !TB)
    gemm_kernel<T, true, false, false><<<grid, NT, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, chunk, NXCD);
  else
    gemm_kernel<T, true, true, false><<<grid, NT, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, chunk, NXCD);
}

template <class T>
void launch_update_t(int m, int n, int k, const T* A, long lda, const T* B, long ldb, T* C, long ldc, bool skip0,
                     hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  const int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
  const int nx = skip0 ? NXCD - 1 : NXCD;
  const int chunk = (tm * tn + nx - 1) / nx;
  gemm_kernel<T, false, false, true><<<dim3(chunk * NXCD), NT, 0, s>>>(m, n, k, T(-1), A, lda, B, ldb, T(1), C, ldc,
                                                                     tm, tn, chunk, nx);
}

}  // namespace

void launch_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, long lda,
                  const double* B, long ldb, double beta, double* C, long ldc, hipStream_t s) {
  launch_gemm_t<double>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
void launch_dgemm_update(int m, int n, int k, const double* A, long lda, const double* B, long ldb, double* C,
                         long ldc, bool skip0, hipStream_t s) {
  launch_update_t<double>(m, n, k, A, lda, B, ldb, C, ldc, skip0, s);
}
void launch_sgemm_update(int m, int n, int k, const float* A, long lda, const float* B, long ldb, float* C, long ldc,
                         bool skip0, hipStream_t s) {
  launch_update_t<float>(m, n, k, A, lda, B, ldb, C, ldc, skip0, s);
}
void launch_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float* A, long lda,
                  const float* B, long ldb, float beta, float* C, long ldc, hipStream_t s) {
  launch_gemm_t<float>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace la
