// gemm_h.hip -- the trailing update C <- C - A*B of an FP32-stored LU with
// operands rounded to FP16 or BF16 and products accumulated in FP32 on the
// half-precision matrix cores (v_mfma_f32_16x16x16_f16 / _bf16).
//
// This is the low-precision factorization of the mixed-precision solver:
// the panel and the triangular solves stay in FP32, the Schur-complement
// updates (SRC/dgetrf.f:212-216, the O(n^3) part) take half-precision
// operands.  There is no reference counterpart; its accuracy contract is the
// refinement's (the solver iterates to the FP64 backward error).
//
// Tiling as the FP64/FP32 kernel (128x128 tile, 2x2 waves of 4x4 MFMA blocks,
// XCD-aware order, XCD 0 optionally left to the panel) with BK = 32: each
// thread converts 16 values of A (one row i, 16 consecutive k) and 16 of B
// (one column j, 16 consecutive k) per k-tile and stores them k-contiguous
// in LDS, where one MFMA operand (4 consecutive k of one row/column) is a
// single 8-byte read.  MFMA layouts: lane l supplies A[l%16][4(l/16)..+3]
// and B[4(l/16)..+3][l%16]; accumulator r of lane l is D[4(l/16)+r][l%16].
// Products are issued as C^T = B^T A^T, so register r of lane l of block
// (i, j) holds C(16i + l%16, 16j + 4(l/16) + r): each C access instruction
// covers 16 consecutive rows of 4 columns (64-byte pieces) instead of
// scattered 4-byte pieces of 16 columns.
#include "la_common.h"

namespace la {
namespace {

constexpr int HBM_ = 128, HBN = 128, HBK = 32, HNT = 256;
constexpr int HGROUP_M = 8;
constexpr int HNXCD = 8;
constexpr int HST = HBK + 8;  // LDS row stride in halves (80 bytes)

typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
typedef short s4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// float -> 16-bit pattern
template <int KIND> __device__ __forceinline__ unsigned short to_half_bits(float x);
template <> __device__ __forceinline__ unsigned short to_half_bits<1>(float x) {  // IEEE FP16, RNE
  const _Float16 h = (_Float16)x;
  return __builtin_bit_cast(unsigned short, h);
}
template <> __device__ __forceinline__ unsigned short to_half_bits<2>(float x) {  // BF16, RNE
  unsigned u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

template <int KIND> __device__ __forceinline__ f4_t mfma_h(unsigned long long a, unsigned long long b, f4_t c);
template <> __device__ __forceinline__ f4_t mfma_h<1>(unsigned long long a, unsigned long long b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(h4_t, a), __builtin_bit_cast(h4_t, b), c, 0, 0, 0);
}
template <> __device__ __forceinline__ f4_t mfma_h<2>(unsigned long long a, unsigned long long b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4_t, a), __builtin_bit_cast(s4_t, b), c, 0, 0,
                                                   0);
}

// FP16 keeps 5 exponent bits: multipliers of L far below 1 would fall into
// its subnormal range.  B here is the block of L (|l| <= 1 under partial
// pivoting), so it is scaled by 2^12 before rounding (still <= 4096 <
// 65504) and the product by 2^-12 after accumulation -- both exact.  BF16
// has FP32's exponent range and needs no scaling.
template <int KIND>
__global__ __launch_bounds__(HNT, 2) void gemm_update_h_kernel(int M, int N, int K, const float* __restrict__ A,
                                                               long lda, const float* __restrict__ B, long ldb,
                                                               float* __restrict__ C, long ldc, int tm, int tn,
                                                               int chunk, int nxcd) {
  constexpr float BSCALE = (KIND == 1) ? 4096.0f : 1.0f;
  __shared__ __align__(16) unsigned short As[2][HBM_ * HST];
  __shared__ __align__(16) unsigned short Bs[2][HBN * HST];

  const int bid = blockIdx.x;
  int xcd = bid % HNXCD;
  if (nxcd < HNXCD) {
    if (xcd < HNXCD - nxcd) return;
    xcd -= HNXCD - nxcd;
  }
  const int lin = xcd * chunk + bid / HNXCD;
  if (lin >= tm * tn) return;
  const int width = HGROUP_M * tn;
  const int group = lin / width;
  const int first_m = group * HGROUP_M;
  const int gsize = min(tm - first_m, HGROUP_M);
  const int tile_m = first_m + (lin % width) % gsize;
  const int tile_n = (lin % width) / gsize;
  const int row0 = tile_m * HBM_;
  const int col0 = tile_n * HBN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0, 0, 0, 0};

  // thread's share of a k-tile: A row ai, k = ak0..ak0+15; B column bj, k = bk0..bk0+15
  const int ai = tid & 127, ak0 = (tid >> 7) * 16;
  const int bj = tid >> 1, bk0 = (tid & 1) * 16;
  float ra[16], rb[16];
  const int nkt = (K + HBK - 1) / HBK;

  auto load_tile = [&](int kt) {
    const int k0 = kt * HBK;
    const int gi = row0 + ai;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int gk = k0 + ak0 + q;
      ra[q] = (gi < M && gk < K) ? A[(long)gi + (long)gk * lda] : 0.0f;
    }
    const int gj = col0 + bj;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int gk = k0 + bk0 + q;
      rb[q] = (gj < N && gk < K) ? B[(long)gk + (long)gj * ldb] : 0.0f;
    }
  };
  auto store_tile = [&](int buf) {
    unsigned pa[8], pb[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      pa[q] = (unsigned)to_half_bits<KIND>(ra[2 * q]) | ((unsigned)to_half_bits<KIND>(ra[2 * q + 1]) << 16);
      pb[q] = (unsigned)to_half_bits<KIND>(BSCALE * rb[2 * q]) |
              ((unsigned)to_half_bits<KIND>(BSCALE * rb[2 * q + 1]) << 16);
    }
    uint4* da = reinterpret_cast<uint4*>(&As[buf][ai * HST + ak0]);
    uint4* db = reinterpret_cast<uint4*>(&Bs[buf][bj * HST + bk0]);
    da[0] = uint4{pa[0], pa[1], pa[2], pa[3]};
    da[1] = uint4{pa[4], pa[5], pa[6], pa[7]};
    db[0] = uint4{pb[0], pb[1], pb[2], pb[3]};
    db[1] = uint4{pb[4], pb[5], pb[6], pb[7]};
  };

  if (nkt > 0) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();

  const int li = lane & 15, lk = lane >> 4;
  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nkt) load_tile(kt + 1);
    const unsigned short* as = As[cur];
    const unsigned short* bs = Bs[cur];
#pragma unroll
    for (int kk = 0; kk < HBK / 16; ++kk) {
      unsigned long long af[4], bf[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        af[b] = *reinterpret_cast<const unsigned long long*>(&as[(wm * 64 + b * 16 + li) * HST + kk * 16 + 4 * lk]);
        bf[b] = *reinterpret_cast<const unsigned long long*>(&bs[(wn * 64 + b * 16 + li) * HST + kk * 16 + 4 * lk]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma_h<KIND>(bf[j], af[i], acc[i][j]);
    }
    if (kt + 1 < nkt) store_tile(cur ^ 1);
    __syncthreads();
  }

  // A 16-row band of C (16 elements per lane) is loaded with clamped,
  // unconditional addresses before its first use, so the band's memory
  // round trips overlap instead of following one another.
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gr = row0 + wm * 64 + i * 16 + li;
    const int grc = min(gr, M - 1);
    float cv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) cv[j][r] = C[(long)grc + (long)min(col0 + wn * 64 + j * 16 + 4 * lk + r, N - 1) * ldc];
    if (gr >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gc = col0 + wn * 64 + j * 16 + 4 * lk + r;
        if (gc < N) C[(long)gr + (long)gc * ldc] = cv[j][r] - acc[i][j][r] * (1.0f / BSCALE);
      }
  }
}

}  // namespace

// kind: 1 = FP16 operands, 2 = BF16 operands (FP32 accumulation either way).
void launch_sgemm_update_half(int kind, int m, int n, int k, const float* A, long lda, const float* B, long ldb,
                              float* C, long ldc, int nskip, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  const int tm = (m + HBM_ - 1) / HBM_, tn = (n + HBN - 1) / HBN;
  const int nx = HNXCD - nskip;
  const int chunk = (tm * tn + nx - 1) / nx;
  if (kind == 2)
    gemm_update_h_kernel<2><<<dim3(chunk * HNXCD), HNT, 0, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tm, tn, chunk, nx);
  else
    gemm_update_h_kernel<1><<<dim3(chunk * HNXCD), HNT, 0, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tm, tn, chunk, nx);
}

}  // namespace la
