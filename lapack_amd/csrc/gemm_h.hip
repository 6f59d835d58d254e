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

#include <cstdlib>

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

// The same update for large updates with the operand feed of the FP32
// direct-to-LDS kernel (update_dma_kernel<float> in gemm.hip): 256 x 128
// tile, 8 waves of 64 x 64, each k stage of 16 moved from HBM straight into
// LDS as FP32 (buffer_load_dwordx4 ... lds), double-buffered, one barrier per
// stage.  The operands are rounded to the half format when read:
//   * B (the block of L): lane (li, lk) reads the four consecutive k
//     4*lk .. 4*lk+3 of column 16*jb + li as one 16-byte LDS read: exactly its
//     four k slots of v_mfma_f32_16x16x16_f16 for the whole stage;
//   * A: four 16-byte reads (k rows 4*lk + q, q = 0..3) give rows 4*li + e,
//     e = 0..3, of the four MFMA blocks (block e holds rows 4*li + e), so
//     component e of the four reads is block e's operand.
// Accumulator r of lane l of block (e, jb) is C(4*(l%16) + e, 16*jb +
// 4*(l/16) + r): the four blocks e of one (jb, r) are four consecutive rows,
// so C is read and written as 16-byte vectors.  As in gemm_update_h_kernel,
// products accumulate from zero and C - acc * 2^-12 (FP16) or C - acc (BF16)
// is stored.  Conditions as update_dma_kernel: A, B, C 16-byte aligned, M and
// the leading dimensions multiples of 4, K of 16 (the launcher checks).
typedef __attribute__((address_space(3))) void h_lds_void;
typedef float f4v_t __attribute__((ext_vector_type(4)));
typedef _Float16 h4v_t __attribute__((ext_vector_type(4)));
typedef __bf16 b4v_t __attribute__((ext_vector_type(4)));
// range never checked (num_records = 2^31 - 1): every offset stays inside
// the operand's own matrix
__device__ __forceinline__ __amdgpu_buffer_rsrc_t hbuf_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0x7fffffff, 0x00020000);
}
// 16 bytes per lane into LDS (a plain function: written inside a kernel
// template the builtin makes the host pass drop the kernel's launch stub)
__device__ __forceinline__ void hlds_dma16(__amdgpu_buffer_rsrc_t r, void* lds, int voff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (h_lds_void*)lds, 16, voff, 0, 0, 0);
}
template <int KIND> __device__ __forceinline__ unsigned long long half4_bits(f4v_t v);
template <> __device__ __forceinline__ unsigned long long half4_bits<1>(f4v_t v) {
  return __builtin_bit_cast(unsigned long long, __builtin_convertvector(v, h4v_t));
}
template <> __device__ __forceinline__ unsigned long long half4_bits<2>(f4v_t v) {
  return __builtin_bit_cast(unsigned long long, __builtin_convertvector(v, b4v_t));
}

template <int KIND>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
void update_dma_h_kernel(int M, int N, int K, const float* __restrict__ A, long lda, const float* __restrict__ B,
                         long ldb, float* __restrict__ C, long ldc, int tm, int tn, int chunk, int nxcd) {
  constexpr float BSCALE = (KIND == 1) ? 4096.0f : 1.0f;
  constexpr int TBM = 256, TBN = 128, KS = 16;
  constexpr int ASZ = KS * TBM, BSZ = TBN * KS;
  __shared__ __align__(16) float As[2][ASZ];
  __shared__ __align__(16) float Bs[2][BSZ];

  const int bid = blockIdx.x;
  int xcd = bid % HNXCD;
  if (nxcd < HNXCD) {
    if (xcd < HNXCD - nxcd) return;
    xcd -= HNXCD - nxcd;
  }
  const int per = gridDim.x / HNXCD;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 3, wn = wave >> 2;
  const int li = lane & 15, lk = lane >> 4;
  const int nkt = K / KS;
#pragma nounroll
  for (int slot = bid / HNXCD; slot < chunk; slot += per) {
    const int lin = __builtin_amdgcn_readfirstlane(xcd * chunk + slot);
    if (lin >= tm * tn) break;
    const int width = HGROUP_M * tn;
    const int group = lin / width;
    const int first_m = group * HGROUP_M;
    const int gsize = min(tm - first_m, HGROUP_M);
    const int row0 = (first_m + (lin % width) % gsize) * TBM;
    const int col0 = ((lin % width) / gsize) * TBN;

    // this wave's loads: k rows wave and wave + 8 of A (256 rows each, four
    // per lane), and 16-column block wave of B (four k per lane); rows past
    // M and columns past N read clamped in-bounds pieces that reach only C
    // elements which are not stored
    int aoff[2];
#pragma unroll
    for (int q = 0; q < 2; ++q)
      aoff[q] = (min(row0 + 4 * lane, M - 4) - row0 + (wave + 8 * q) * (int)lda) * (int)sizeof(float);
    const int boff = (4 * lk + (min(col0 + 16 * wave + li, N - 1) - col0) * (int)ldb) * (int)sizeof(float);
    auto issue = [&](int kt, int buf) {
      const long k0 = (long)kt * KS;
      const __amdgpu_buffer_rsrc_t ra = hbuf_rsrc(A + row0 + k0 * lda);
      const __amdgpu_buffer_rsrc_t rb = hbuf_rsrc(B + k0 + (long)col0 * ldb);
#pragma unroll
      for (int q = 0; q < 2; ++q) hlds_dma16(ra, &As[buf][(wave + 8 * q) * TBM], aoff[q]);
      hlds_dma16(rb, &Bs[buf][wave * 256], boff);
    };
    issue(0, 0);
    f4_t acc[4][4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int jb = 0; jb < 4; ++jb) acc[e][jb] = f4_t{0, 0, 0, 0};
    for (int kt = 0; kt < nkt; ++kt) {
      const int cur = kt & 1;
      __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's loads of stage kt are in LDS
      __syncthreads();
      f4v_t af[4], bf[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) af[q] = *(const f4v_t*)&As[cur][(4 * lk + q) * TBM + wm * 64 + 4 * li];
#pragma unroll
      for (int jb = 0; jb < 4; ++jb) bf[jb] = *(const f4v_t*)&Bs[cur][(wn * 4 + jb) * 256 + lane * 4];
      if (kt + 1 < nkt) issue(kt + 1, cur ^ 1);
      unsigned long long ah[4], bh[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) ah[e] = half4_bits<KIND>(f4v_t{af[0][e], af[1][e], af[2][e], af[3][e]});
#pragma unroll
      for (int jb = 0; jb < 4; ++jb) bh[jb] = half4_bits<KIND>(bf[jb] * BSCALE);
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) acc[e][jb] = mfma_h<KIND>(bh[jb], ah[e], acc[e][jb]);
    }
    // every wave is done with the last stage's buffer before the next tile's
    // stage 0 is loaded (the same buffer when K/16 is odd)
    __syncthreads();
    const int rw = row0 + wm * 64 + 4 * li;
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
      f4v_t cv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        cv[r] = *(const f4v_t*)(C + (long)min(col0 + wn * 64 + 16 * jb + 4 * lk + r, N - 1) * ldc + min(rw, M - 4));
      if (rw >= M) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gc = col0 + wn * 64 + 16 * jb + 4 * lk + r;
        if (gc >= N) continue;
        f4v_t v;
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = cv[r][e] - acc[e][jb][r] * (1.0f / BSCALE);
        *(f4v_t*)(C + (long)gc * ldc + rw) = v;
      }
    }
  }
}

}  // namespace

// kind: 1 = FP16 operands, 2 = BF16 operands (FP32 accumulation either way).
void launch_sgemm_update_half(int kind, int m, int n, int k, const float* A, long lda, const float* B, long ldb,
                              float* C, long ldc, int nskip, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  const int nx = HNXCD - nskip;
  // large updates on the direct-to-LDS kernel where it fits (the thresholds
  // and switches of the FP32 update: LAPACK_AMD_UPD_DMA, _UPD_BIG_MIN,
  // _UPD_DMA_OCC); beside a panel one workgroup per CU (40 KB of dynamic LDS)
  {
    const char* e = std::getenv("LAPACK_AMD_UPD_DMA");
    const bool dma = !(e && *e && std::atoi(e) == 0);
    const char* eb = std::getenv("LAPACK_AMD_UPD_BIG_MIN");
    const long big_min = (eb && *eb) ? std::atol(eb) : 512;
    const char* eo = std::getenv("LAPACK_AMD_UPD_DMA_OCC");
    const int occ = (eo && *eo) ? std::atoi(eo) : 1;
    const int tmb = (m + 255) / 256, tnb = (n + 127) / 128;
    const unsigned long al = (unsigned long)A | (unsigned long)B | (unsigned long)C;
    const bool fits = (al & 15) == 0 && m % 4 == 0 && k % 16 == 0 && lda % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 &&
                      (16 * lda + 256) * 4L < (1L << 31) && (128 * ldb + 16) * 4L < (1L << 31);
    if (dma && fits && (long)tmb * tnb >= big_min) {
      const int chunk = (tmb * tnb + nx - 1) / nx;
      const size_t dyn = occ == 1 ? 40 * 1024 : 0;
      if (kind == 2)
        update_dma_h_kernel<2><<<dim3(chunk * HNXCD), 512, dyn, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tmb, tnb, chunk, nx);
      else
        update_dma_h_kernel<1><<<dim3(chunk * HNXCD), 512, dyn, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tmb, tnb, chunk, nx);
      return;
    }
  }
  const int tm = (m + HBM_ - 1) / HBM_, tn = (n + HBN - 1) / HBN;
  const int chunk = (tm * tn + nx - 1) / nx;
  if (kind == 2)
    gemm_update_h_kernel<2><<<dim3(chunk * HNXCD), HNT, 0, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tm, tn, chunk, nx);
  else
    gemm_update_h_kernel<1><<<dim3(chunk * HNXCD), HNT, 0, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tm, tn, chunk, nx);
}

}  // namespace la
