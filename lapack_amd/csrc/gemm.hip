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
// Tiling (general kernel): a 256-thread workgroup (4 waves, 2x2) owns a
// 128x128 tile of C;
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

#include <algorithm>
#include <cstdlib>

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

// Products are issued as C^T = B^T A^T (the B fragment in the MFMA's A
// slot), so block (i, j) of a wave's 64x64 C tile holds, in register r of
// lane l, C(16i + l%16, 16j + Mfma<T>::row(l/16, r)): for each register the
// 16 lanes of a group cover 16 consecutive rows of one column, whole
// 128-byte lines for FP64 and half lines for FP32.  With the operands the
// other way round every C access instruction touched 16 lines per wave in
// 32-byte (FP64) or 4-byte (FP32) pieces; measured on MI355X at m = n = 16128,
// k = 256 the C traffic then ran at 2.0-4.1 TB/s against 5.1 for the
// coalesced form, 41.1 -> 46.9 TFLOP/s (FP64) and 61.2 -> 84.9 (FP32).
//
// The update forms read C into the accumulators before the k loop
// (acc = C, then acc += (-A) B, the order of BLAS/SRC/dgemm.f:265-276 with
// alpha = -1), so the epilogue only stores.  Rows and columns past the edge
// of C read a clamped in-bounds element and are never stored.
template <class T, class Acc>
__device__ __forceinline__ void preload_c(Acc (&acc)[4][4], const T* __restrict__ C, long ldc, int M, int N,
                                          int r0, int c0, int li, int lk) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const T* cp = C + (long)min(c0 + j * 16 + Mfma<T>::row(lk, r), N - 1) * ldc;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i][j][r] = cp[min(r0 + i * 16 + li, M - 1)];
    }
}

// Stores scale * acc for the in-bounds elements (no loads).
template <class T, class Acc>
__device__ __forceinline__ void store_c(const Acc (&acc)[4][4], T* __restrict__ C, long ldc, int M, int N, int r0,
                                        int c0, int li, int lk, T scale) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gc = c0 + j * 16 + Mfma<T>::row(lk, r);
      if (gc >= N) continue;
      T* cp = C + (long)gc * ldc;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gr = r0 + i * 16 + li;
        if (gr < M) cp[gr] = scale == T(1) ? acc[i][j][r] : scale * acc[i][j][r];
      }
    }
}

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

  const int li = lane & 15;  // row (A) / column (B) inside a 16-block
  const int lk = lane >> 4;  // k inside a 4-step
  acc_t acc[4][4];

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
      As[buf][k * AST + i] = UPD ? -ra[q] : ra[q];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int j, k;
      if (!TB) { k = tid & 15;  j = (tid >> 4) + 16 * q; }
      else     { j = tid & 127; k = (tid >> 7) + 2 * q; }
      Bs[buf][j * BST + k] = rb[q];
    }
  };

  // first A/B tile in flight before the C preload, so that its LDS store
  // waits for those loads only
  if (nkt > 0) load_tile(0);
  if (UPD)
    preload_c<T>(acc, C, ldc, M, N, row0 + wm * 64, col0 + wn * 64, li, lk);
  else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  }
  if (nkt > 0) store_tile(0);
  __syncthreads();

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
        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(bf[j], af[i], acc[i][j]);
    }
    if (kt + 1 < nkt) store_tile(cur ^ 1);
    __syncthreads();
  }

  // Epilogue (element layout as described at preload_c).
  if (UPD || beta == T(0)) {
    store_c<T>(acc, C, ldc, M, N, row0 + wm * 64, col0 + wn * 64, li, lk, UPD ? T(1) : alpha);
    return;
  }
  // beta != 0: a 16-row band of C at a time is read with every load in
  // flight before the first use (clamped, unconditional addresses), then
  // combined and stored.
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gr = row0 + wm * 64 + i * 16 + li;
    T cv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gc = min(col0 + wn * 64 + j * 16 + Mfma<T>::row(lk, r), N - 1);
        cv[j][r] = C[(long)min(gr, M - 1) + (long)gc * ldc];
      }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gc = col0 + wn * 64 + j * 16 + Mfma<T>::row(lk, r);
        if (gr < M && gc < N) C[(long)gr + (long)gc * ldc] = alpha * acc[i][j][r] + beta * cv[j][r];
      }
  }
}

// The trailing update C <- C - A*B with A, B, C untransposed, for large
// updates: a 512-thread workgroup (8 waves, WM x WN) owns a (WM*64) x (WN*64)
// tile, each wave 4x4 MFMA blocks as above.  Eight waves per CU (two per
// SIMD) hide the LDS and MFMA latencies that the 4-wave kernel above leaves
// exposed; measured on MI355X at m = n = 16128, k = 256: 38.6 TFLOP/s
// against 30.1 for the 128x128 tile.
template <class T, int WM, int WN>
__global__ __launch_bounds__(WM * WN * 64) void update_big_kernel(int M, int N, int K, const T* __restrict__ A,
                                                                  long lda, const T* __restrict__ B, long ldb,
                                                                  T* __restrict__ C, long ldc, int tm, int tn,
                                                                  int chunk, int nxcd) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int NTB = WM * WN * 64;
  constexpr int TBM = WM * 64, TBN = WN * 64;
  constexpr int AST = TBM + Mfma<T>::APAD;
  constexpr int BST = BK + 1;
  constexpr int LA = TBM * BK / NTB, LB = TBN * BK / NTB;
  __shared__ T As[2][BK * AST];
  __shared__ T Bs[2][TBN * BST];

  const int bid = blockIdx.x;
  int xcd = bid % NXCD;
  if (nxcd < NXCD) {
    if (xcd < NXCD - nxcd) return;
    xcd -= NXCD - nxcd;
  }
  // the XCD's blocks walk its chunk of tiles (one pass when the grid holds
  // a block per tile; fewer, persistent blocks leave CUs free)
  const int per = gridDim.x / NXCD;
#pragma nounroll
  for (int slot = bid / NXCD; slot < chunk; slot += per) {
  const int lin = __builtin_amdgcn_readfirstlane(xcd * chunk + slot);
  if (lin >= tm * tn) break;
  const int width = GROUP_M * tn;
  const int group = lin / width;
  const int first_m = group * GROUP_M;
  const int gsize = min(tm - first_m, GROUP_M);
  const int tile_m = first_m + (lin % width) % gsize;
  const int tile_n = (lin % width) / gsize;
  const int row0 = tile_m * TBM;
  const int col0 = tile_n * TBN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM;
  const int wn = wave / WM;

  const int li = lane & 15, lk = lane >> 4;
  acc_t acc[4][4];

  T ra[LA], rb[LB];
  const int nkt = (K + BK - 1) / BK;
  // thread's fixed row i (A) / k (B); the q-th element is one uniform
  // stride further, so a single pointer per operand stays live
  static_assert(NTB % TBM == 0 && NTB % BK == 0, "tile shape");
  const int ai = tid % TBM;
  const int ak = __builtin_amdgcn_readfirstlane(tid / TBM);  // uniform per wave (TBM is a multiple of 64)
  const int bk = tid % BK, bj = tid / BK;
  auto load_tile = [&](int kt) {
    const int k0 = kt * BK;
    const bool arow = row0 + ai < M;
    const T* ap = A + (long)(row0 + ai) + (long)(k0 + ak) * lda;
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int gk = k0 + ak + q * (NTB / TBM);
      ra[q] = (arow && gk < K) ? ap[(long)(q * (NTB / TBM)) * lda] : T(0);
    }
    const bool bcol = k0 + bk < K;
    const T* bp = B + (long)(k0 + bk) + (long)(col0 + bj) * ldb;
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int gj = col0 + bj + q * (NTB / BK);
      rb[q] = (bcol && gj < N) ? bp[(long)(q * (NTB / BK)) * ldb] : T(0);
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int e = tid + q * NTB;
      As[buf][(e / TBM) * AST + e % TBM] = -ra[q];
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int e = tid + q * NTB;
      Bs[buf][(e / BK) * BST + e % BK] = rb[q];
    }
  };
  if (nkt > 0) load_tile(0);
  preload_c<T>(acc, C, ldc, M, N, row0 + wm * 64, col0 + wn * 64, li, lk);
  if (nkt > 0) store_tile(0);
  __syncthreads();
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
        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(bf[j], af[i], acc[i][j]);
    }
    if (kt + 1 < nkt) store_tile(cur ^ 1);
    __syncthreads();
  }
  store_c<T>(acc, C, ldc, M, N, row0 + wm * 64, col0 + wn * 64, li, lk, T(1));
  }
}

// The trailing update C <- C - A*B with direct global -> LDS loads
// (buffer_load_dwordx4 ... lds: each lane moves 16 bytes into LDS, the wave's
// 64 pieces land contiguously at a wave-uniform LDS address), for A, B, C
// 16-byte aligned, M and the leading dimensions multiples of the 16-byte
// vector and K a multiple of 16 (the launcher checks; other shapes keep
// update_big_kernel).  Tile 256 x 128, 8 waves of 64 x 64 as above.
//
// No staging registers: a k stage of 16 goes straight to LDS, and every
// fragment a wave needs from it is one 16-byte LDS read:
//   * lane (li, lk) takes k = 4*lk + kk in MFMA step kk (any assignment of
//     the 16 k of a stage to (lk, kk) is a valid product), so the four k of
//     one lane are consecutive in B's column: B's stage is stored as 16-
//     column blocks, piece p of lane L at block*256 + p*64*VE + L*VE, one
//     contiguous 1 KB wave read per block and piece;
//   * a lane's A fragments are VE consecutive rows of one k (rows VE*li+e of
//     the half h), one read giving VE of the wave's four 16-row MFMA blocks;
//     block b = h*VE + e then holds rows 16*VE*h + VE*li + e, so the
//     accumulators of the blocks e = 0..VE-1 are VE consecutive rows of one
//     column of C: C is read and written as 16-byte vectors.
// Each stage's fragments are all read before the next stage's loads are
// issued into the other buffer, so those loads get the whole MFMA phase to
// arrive; one barrier per stage (after this wave's loads are complete)
// publishes the stage.  A stage holds NSUB sub-stages of 16 k (NSUB = 2:
// FP32 beside a panel, which also keeps C aside and adds it after the
// products; see the kernel).
template <class T> struct V16;
template <> struct V16<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct V16<double> { typedef double type __attribute__((ext_vector_type(2))); };
typedef __attribute__((address_space(3))) void la_lds_void;
// Buffer-resource access (MUBUF): a uniform 64-bit base in scalar registers
// plus a 32-bit per-lane byte offset and a uniform byte offset, so a tile's
// loads and stores keep one offset register per lane instead of one 64-bit
// address per element.  The range is never checked (num_records = 2^31-1;
// every offset used stays inside the tile's own matrix).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0x7fffffff, 0x00020000);
}
// 16 bytes per lane from buffer r at byte offset voff into LDS: the wave's
// 64 pieces land contiguously from the wave-uniform address lds.  (A plain
// function: with the builtin written inside a kernel template the host
// compilation pass drops that kernel's launch stub.)
__device__ __forceinline__ void lds_dma16(__amdgpu_buffer_rsrc_t r, void* lds, int voff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (la_lds_void*)lds, 16, voff, 0, 0, 0);
}

// NSUB: k sub-stages of 16 per LDS stage.  With 2 (FP32 only: 96 KB of LDS,
// one workgroup per CU) each stage's loads have two sub-stages of MFMAs to
// arrive, and there is one barrier per 32 k.
template <class T, int NSUB>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu((sizeof(T) == 4 && NSUB == 1) ? 4 : 2)))
void update_dma_kernel(int M, int N, int K, const T* __restrict__ A, long lda, const T* __restrict__ B, long ldb,
                       T* __restrict__ C, long ldc, int tm, int tn, int chunk, int nxcd) {
  typedef typename Mfma<T>::acc_t acc_t;
  typedef typename V16<T>::type v_t;
  constexpr int VE = 16 / (int)sizeof(T);  // elements per 16-byte piece
  constexpr int NP = 4 / VE;               // pieces per lane and k stage (FP32 1, FP64 2)
  constexpr int TBM = 256, TBN = 128, KS = 16;
  constexpr int ASZ = KS * TBM, BSZ = TBN * KS;  // elements of one stage
  constexpr int CPR = TBM * (int)sizeof(T) / 1024;  // wave loads per k row of A
  constexpr int NLA = ASZ * (int)sizeof(T) / 1024 / 8, NLB = BSZ * (int)sizeof(T) / 1024 / 8;  // per wave
  __shared__ __align__(16) T As[2][NSUB * ASZ];
  __shared__ __align__(16) T Bs[2][NSUB * BSZ];

  const int bid = blockIdx.x;
  int xcd = bid % NXCD;
  if (nxcd < NXCD) {
    if (xcd < NXCD - nxcd) return;
    xcd -= NXCD - nxcd;
  }
  const int per = gridDim.x / NXCD;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 3, wn = wave >> 2;
  const int li = lane & 15, lk = lane >> 4;
  const int nkt = K / (KS * NSUB);
#pragma nounroll
  for (int slot = bid / NXCD; slot < chunk; slot += per) {
    const int lin = __builtin_amdgcn_readfirstlane(xcd * chunk + slot);
    if (lin >= tm * tn) break;
    const int width = GROUP_M * tn;
    const int group = lin / width;
    const int first_m = group * GROUP_M;
    const int gsize = min(tm - first_m, GROUP_M);
    const int row0 = (first_m + (lin % width) % gsize) * TBM;
    const int col0 = ((lin % width) / gsize) * TBN;

    // this wave's loads of stage kt (sub-stage su: k from (kt*NSUB+su)*16) into buffer buf, as buffer loads from
    // the stage's corner (uniform) plus a byte offset per lane fixed for the
    // tile (below 2^31: the launcher checks 16*lda and 128*ldb): rows past M
    // and columns past N read clamped in-bounds pieces (their products reach
    // only rows and columns of C that are not stored)
    int aoff[NLA], boff[NLB];
#pragma unroll
    for (int q = 0; q < NLA; ++q) {
      const int a = wave + 8 * q, ka = a / CPR, c = a % CPR;
      aoff[q] = (min(row0 + c * 64 * VE + VE * lane, M - VE) - row0 + ka * (int)lda) * (int)sizeof(T);
    }
#pragma unroll
    for (int q = 0; q < NLB; ++q) {
      const int b = wave + 8 * q, jb = b / NP, p = b % NP;
      boff[q] = (4 * lk + VE * p + (min(col0 + 16 * jb + li, N - 1) - col0) * (int)ldb) * (int)sizeof(T);
    }
    auto issue = [&](int kt, int buf) {
#pragma unroll
      for (int su = 0; su < NSUB; ++su) {
        const long k0 = ((long)kt * NSUB + su) * KS;
        const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A + row0 + k0 * lda);
        const __amdgpu_buffer_rsrc_t rb = buf_rsrc(B + k0 + (long)col0 * ldb);
#pragma unroll
        for (int q = 0; q < NLA; ++q) {
          const int a = wave + 8 * q, ka = a / CPR, c = a % CPR;
          lds_dma16(ra, &As[buf][su * ASZ + ka * TBM + c * 64 * VE], aoff[q]);
        }
#pragma unroll
        for (int q = 0; q < NLB; ++q) {
          const int b = wave + 8 * q, jb = b / NP, p = b % NP;
          lds_dma16(rb, &Bs[buf][su * BSZ + jb * 256 + p * 64 * VE], boff[q]);
        }
      }
    };

    // stage 0 first: the reads of C go out behind it.  With stages of 32 k
    // (one workgroup per CU, nothing else to cover a tile's start) C is kept
    // aside and added after the products (LATE_C: C + sum instead of the sum
    // started from C), so its reads overlap the first stage instead of
    // preceding it.
    constexpr bool LATE_C = NSUB == 2;
    constexpr int NCL = 4 * 4 * NP;  // this lane's reads of C
    issue(0, 0);
    __builtin_amdgcn_sched_barrier(0);
    acc_t acc[4][4];
    v_t cpre[4][4][NP];
    const int rw = row0 + wm * 64 + VE * li;  // + 16*VE*h
    const int cw = col0 + wn * 64;            // + 16*jb + row(lk, r)
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const T* cp = C + (long)min(cw + 16 * jb + Mfma<T>::row(lk, r), N - 1) * ldc;
#pragma unroll
        for (int h = 0; h < NP; ++h) {
          const v_t v = *(const v_t*)(cp + min(rw + 16 * VE * h, M - VE));
          if (LATE_C) {
            cpre[jb][r][h] = v;
          } else {
#pragma unroll
            for (int e = 0; e < VE; ++e) acc[h * VE + e][jb][r] = v[e];
          }
        }
      }
    if (LATE_C) {
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) acc[b][jb] = acc_t{};
    }
    static_assert(!LATE_C || NCL == 16, "vmcnt(16) below counts the reads of C");
    for (int kt = 0; kt < nkt; ++kt) {
      const int cur = kt & 1;
      // this wave's loads of stage kt are in LDS: vmcnt(0), or at the first
      // stage with LATE_C vmcnt(16) (the reads of C, issued after it, may
      // still be in flight)
      if (LATE_C && kt == 0) __builtin_amdgcn_s_waitcnt(0x4f70);
      else __builtin_amdgcn_s_waitcnt(0x0f70);
      __syncthreads();
      v_t af[NSUB][4][NP], bf[NSUB][4][NP];
#pragma unroll
      for (int su = 0; su < NSUB; ++su) {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
          for (int h = 0; h < NP; ++h)
            af[su][kk][h] = *(const v_t*)&As[cur][su * ASZ + (4 * lk + kk) * TBM + wm * 64 + h * 16 * VE + VE * li];
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
#pragma unroll
          for (int p = 0; p < NP; ++p)
            bf[su][jb][p] = *(const v_t*)&Bs[cur][su * BSZ + (wn * 4 + jb) * 256 + p * 64 * VE + lane * VE];
      }
      if (kt + 1 < nkt) issue(kt + 1, cur ^ 1);
#pragma unroll
      for (int su = 0; su < NSUB; ++su)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
          for (int h = 0; h < NP; ++h)
#pragma unroll
            for (int e = 0; e < VE; ++e)
#pragma unroll
              for (int jb = 0; jb < 4; ++jb)
                acc[h * VE + e][jb] =
                    Mfma<T>::op(-bf[su][jb][kk / VE][kk % VE], af[su][kk][h][e], acc[h * VE + e][jb]);
    }
    // every wave is done with the last stage's buffer before the next tile's
    // stage 0 is loaded (the same buffer when K/16 is odd)
    __syncthreads();
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gc = cw + 16 * jb + Mfma<T>::row(lk, r);
        if (gc >= N) continue;
        T* cp = C + (long)gc * ldc;
#pragma unroll
        for (int h = 0; h < NP; ++h) {
          const int gr = rw + 16 * VE * h;
          if (gr >= M) continue;
          v_t v;
#pragma unroll
          for (int e = 0; e < VE; ++e) v[e] = LATE_C ? cpre[jb][r][h][e] + acc[h * VE + e][jb][r] : acc[h * VE + e][jb][r];
          *(v_t*)(cp + gr) = v;
        }
      }
  }
}

typedef int la_i2 __attribute__((ext_vector_type(2)));
template <class T> __device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <> __device__ __forceinline__ double buf_ld<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <> __device__ __forceinline__ float buf_ld<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void buf_st(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(la_i2, v), r, voff, soff, 0);
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, voff, soff, 0);
}

// The trailing update C <- C - A*B, persistent and software-pipelined
// across tiles, for M and N multiples of 128 (the launcher gives any edge
// strips to gemm_kernel).  A 512-thread workgroup (8 waves, 4 x 2) owns a
// 128 x 128 tile at a time, each wave 32 x 64 = 2 x 4 MFMA blocks (32 FP64
// accumulators per lane), which leaves registers for a second set: while a
// tile's k loop runs, the C elements of the workgroup's next tile are read
// into that set, one block per k step, issued once the step's A/B slice has
// arrived, so that they have a whole step to arrive before the next step's
// wait for its own slice covers them.  In the tile's last k step the
// A/B loads fetch the next tile's first k slice.  A tile's end is then its C
// stores and a register copy; no read of C or of the first slice is waited
// for between tiles.  BK = 32 per LDS stage (137 KB of LDS for FP64, one
// workgroup per CU).  Every global access is a buffer access from a uniform
// base (the tile's corner, or the k slice's) plus a per-lane and a uniform
// byte offset, whose sum stays below 2^31 while ld * sizeof(T) < 2^24 (the
// launcher checks this).
template <class T>
__global__ __launch_bounds__(512) void update_pipe_kernel(int K, const T* __restrict__ A, long lda,
                                                         const T* __restrict__ B, long ldb, T* __restrict__ C,
                                                         long ldc, int tm, int tn, int chunk, int nxcd) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int NTB = 512, TBM = 128, TBN = 128, PBK = 32;
  constexpr int AST = TBM + Mfma<T>::APAD;
  constexpr int BST = PBK + 1;
  constexpr int LA = TBM * PBK / NTB, LB = TBN * PBK / NTB;
  constexpr int NG = 8;  // accumulator blocks per wave = C read groups per tile
  constexpr int ES = sizeof(T);
  __shared__ T As[2][PBK * AST];
  __shared__ T Bs[2][TBN * BST];

  const int bid = blockIdx.x;
  int xcd = bid % NXCD;
  if (nxcd < NXCD) {
    if (xcd < NXCD - nxcd) return;
    xcd -= NXCD - nxcd;
  }
  const int per = gridDim.x / NXCD;
  const int ntiles = tm * tn;
  int slot = bid / NXCD;
  if (slot >= chunk || xcd * chunk + slot >= ntiles) return;
  auto tile_origin = [&](int s, int& r0, int& c0) {
    const int lin = xcd * chunk + s;
    const int width = GROUP_M * tn;
    const int group = lin / width;
    const int first_m = group * GROUP_M;
    const int gsize = min(tm - first_m, GROUP_M);
    r0 = (first_m + (lin % width) % gsize) * TBM;
    c0 = ((lin % width) / gsize) * TBN;
  };

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 3;   // 32-row band of the tile
  const int wn = wave >> 2;  // 64-column band
  const int li = lane & 15, lk = lane >> 4;

  acc_t acc[2][4], cb[2][4];
  T ra[LA], rb[LB];
  const int nkt = (K + PBK - 1) / PBK;
  // A slice: thread (ai, ak) loads row ai, k = ak + 4q (ak uniform per wave)
  const int ai = tid % TBM;
  const int ak = wave / 2;
  const int va = ai * ES;
  // B slice: thread (bk, bj) loads k = bk of columns bj + 16q
  const int bk = tid % PBK, bj = tid / PBK;
  const int vb = (int)((bk + bj * ldb) * ES);
  auto load_tile = [&](int r0, int c0, int kt) {
    const int k0 = kt * PBK;
    const __amdgpu_buffer_rsrc_t ra_ = buf_rsrc(A + (long)r0 + (long)k0 * lda);
    const __amdgpu_buffer_rsrc_t rb_ = buf_rsrc(B + (long)k0 + (long)c0 * ldb);
    if (k0 + PBK <= K) {
#pragma unroll
      for (int q = 0; q < LA; ++q) ra[q] = buf_ld<T>(ra_, va, (int)((ak + 4 * q) * lda * ES));
#pragma unroll
      for (int q = 0; q < LB; ++q) rb[q] = buf_ld<T>(rb_, vb, (int)(16 * q * ldb * ES));
    } else {
      // the last, partial k slice: k past K reads as zero
#pragma unroll
      for (int q = 0; q < LA; ++q)
        ra[q] = (k0 + ak + 4 * q < K) ? buf_ld<T>(ra_, va, (int)((ak + 4 * q) * lda * ES)) : T(0);
      const bool kin = k0 + bk < K;
      const int vbk = (int)((min(bk, K - 1 - k0) + bj * ldb) * ES);
#pragma unroll
      for (int q = 0; q < LB; ++q) {
        const T v = buf_ld<T>(rb_, vbk, (int)(16 * q * ldb * ES));
        rb[q] = kin ? v : T(0);
      }
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int q = 0; q < LA; ++q) As[buf][(ak + 4 * q) * AST + ai] = -ra[q];
#pragma unroll
    for (int q = 0; q < LB; ++q) Bs[buf][(bj + 16 * q) * BST + bk] = rb[q];
  };
  // C elements of block (i, j) for this lane: register r holds row
  // r0 + wm*32 + 16i + li, column c0 + wn*64 + 16j + row(lk, r): one
  // per-lane offset, the rest uniform
  const int vc = (int)((wm * 32 + li + (long)(wn * 64 + Mfma<T>::row(lk, 0)) * ldc) * ES);
  auto c_soff = [&](int i, int j, int r) { return (int)((i * 16 + (long)(j * 16 + Mfma<T>::row(0, r)) * ldc) * ES); };
  auto read_c = [&](acc_t& dst, int r0, int c0, int i, int j) {
    const __amdgpu_buffer_rsrc_t rc = buf_rsrc(C + (long)r0 + (long)c0 * ldc);
#pragma unroll
    for (int r = 0; r < 4; ++r) dst[r] = buf_ld<T>(rc, vc, c_soff(i, j, r));
  };

  int row0, col0;
  tile_origin(slot, row0, col0);
  load_tile(row0, col0, 0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) read_c(acc[i][j], row0, col0, i, j);
  store_tile(0);
  // the first tile's C has arrived before the loop: otherwise the
  // compiler's count of outstanding loads at the loop head keeps the
  // prologue's, and every k step waits for all of its loads (the next
  // slice included) before its first MFMA
  wait_vm();
  __syncthreads();
  int it = 0;
#pragma nounroll
  for (;;) {
    const int nslot = slot + per;
    const bool more = nslot < chunk && xcd * chunk + nslot < ntiles;
    int row1 = 0, col1 = 0;
    if (more) tile_origin(nslot, row1, col1);
#pragma nounroll
    for (int kt = 0; kt < nkt; ++kt, ++it) {
      const int cur = it & 1;
      const bool nxt = kt + 1 < nkt || more;
      if (kt + 1 < nkt) load_tile(row0, col0, kt + 1);
      else if (more) load_tile(row1, col1, 0);
      const T* as = As[cur];
      const T* bs = Bs[cur];
      // operand fragments one 4-deep k step ahead of the MFMAs that use
      // them; the scheduling barrier keeps the compiler from hoisting every
      // step's LDS reads to the top (which costs 12 registers per step)
      T af[2][2], bf[2][4];
      auto frag = [&](int kk, int u) {
#pragma unroll
        for (int b = 0; b < 2; ++b) af[u][b] = as[(kk * 4 + lk) * AST + wm * 32 + b * 16 + li];
#pragma unroll
        for (int b = 0; b < 4; ++b) bf[u][b] = bs[(wn * 64 + b * 16 + li) * BST + kk * 4 + lk];
      };
      frag(0, 0);
#pragma unroll
      for (int kk = 0; kk < PBK / 4; ++kk) {
        if (kk + 1 < PBK / 4) frag(kk + 1, (kk + 1) & 1);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(bf[kk & 1][j], af[kk & 1][i], acc[i][j]);
        __builtin_amdgcn_sched_barrier(0);
      }
      if (nxt) store_tile(cur ^ 1);
      // the next tile's C, group g read at step g * nkt / NG, after this
      // step's A/B slice has been waited for: the next step's wait for its
      // own (younger) slice then covers these, a whole step later
      if (more) {
#pragma unroll
        for (int g = 0; g < NG; ++g)
          if ((g * nkt) / NG == kt) read_c(cb[g / 4][g % 4], row1, col1, g / 4, g % 4);
      }
      __syncthreads();
    }
    // this tile's results (C was read into the accumulators)
    {
      const __amdgpu_buffer_rsrc_t rc = buf_rsrc(C + (long)row0 + (long)col0 * ldc);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) buf_st(acc[i][j][r], rc, vc, c_soff(i, j, r));
    }
    if (!more) break;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = cb[i][j];
    slot = nslot;
    row0 = row1;
    col0 = col1;
  }
}

// Small products (C = alpha*A*B + beta*C, A and B untransposed, few 128x128
// tiles: the DTRSM product of the next panel's 256 columns is 256x256x256,
// four such tiles): 64x64 tiles, so that 16 workgroups share the work.  A
// 256-thread workgroup = 2x2 waves of 2x2 MFMA blocks, BK = 16, double-
// buffered LDS, products issued as C^T = B^T A^T as in gemm_kernel.
// UPD: the trailing-update instance (alpha = -1, beta = 1), a symbol of its
// own so that traces and counters tell it from the DTRSM products.
template <class T, bool UPD = false>
__global__ __launch_bounds__(256) void gemm64_kernel(int M, int N, int K, T alpha, const T* __restrict__ A,
                                                     long lda, const T* __restrict__ B, long ldb, T beta,
                                                     T* __restrict__ C, long ldc, int tm) {
  if constexpr (UPD) {
    alpha = T(-1);
    beta = T(1);
  }
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int TS = 64;
  constexpr int AST = TS + Mfma<T>::APAD;
  constexpr int BST = BK + 1;
  __shared__ T As[2][BK * AST];
  __shared__ T Bs[2][TS * BST];
  const int row0 = (blockIdx.x % tm) * TS;
  const int col0 = (blockIdx.x / tm) * TS;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lk = lane >> 4;
  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  T ra[4], rb[4];
  const int nkt = (K + BK - 1) / BK;
  auto load_tile = [&](int kt) {
    const int k0 = kt * BK;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = tid & 63, kk = (tid >> 6) + 4 * q;
      const int gi = row0 + i, gk = k0 + kk;
      ra[q] = (gi < M && gk < K) ? A[(long)gi + (long)gk * lda] : T(0);
      const int bk = tid & 15, j = (tid >> 4) + 16 * q;
      const int gj = col0 + j, gk2 = k0 + bk;
      rb[q] = (gj < N && gk2 < K) ? B[(long)gk2 + (long)gj * ldb] : T(0);
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[buf][((tid >> 6) + 4 * q) * AST + (tid & 63)] = ra[q];
      Bs[buf][((tid >> 4) + 16 * q) * BST + (tid & 15)] = rb[q];
    }
  };
  if (nkt > 0) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nkt) load_tile(kt + 1);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      T af[2], bf[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        af[b] = As[cur][(kk * 4 + lk) * AST + wm * 32 + b * 16 + li];
        bf[b] = Bs[cur][(wn * 32 + b * 16 + li) * BST + kk * 4 + lk];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::op(bf[j], af[i], acc[i][j]);
    }
    if (kt + 1 < nkt) store_tile(cur ^ 1);
    __syncthreads();
  }
  // register r of lane l of block (i, j): C(16i + l%16, 16j + row(l/16, r));
  // with beta != 0 the 16 elements of C a lane owns are read first, all in
  // flight together (clamped addresses), then combined and stored
  T cv[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = min(row0 + wm * 32 + i * 16 + li, M - 1);
        const int gc = min(col0 + wn * 32 + j * 16 + Mfma<T>::row(lk, r), N - 1);
        cv[i][j][r] = beta == T(0) ? T(0) : C[(long)gr + (long)gc * ldc];
      }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int gr = row0 + wm * 32 + i * 16 + li;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gc = col0 + wn * 32 + j * 16 + Mfma<T>::row(lk, r);
        if (gr < M && gc < N)
          C[(long)gr + (long)gc * ldc] = (beta == T(0)) ? alpha * acc[i][j][r] : alpha * acc[i][j][r] + beta * cv[i][j][r];
      }
  }
}

template <class T>
void launch_update_t(int m, int n, int k, const T* A, long lda, const T* B, long ldb, T* C, long ldc, int nskip,
                     hipStream_t s, bool alone = false);

template <class T>
void launch_gemm_t(char ta, char tb, int m, int n, int k, T alpha, const T* A, long lda, const T* B,
                   long ldb, T beta, T* C, long ldc, hipStream_t s, int nskip) {
  if (m <= 0 || n <= 0) return;
  // BLAS quick return (BLAS/SRC/dgemm.f:298-299).
  if ((alpha == T(0) || k <= 0) && beta == T(1)) return;
  const int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
  const int tiles = tm * tn;
  const int nx = NXCD - nskip;
  const int chunk = (tiles + nx - 1) / nx;
  const dim3 grid(chunk * NXCD);
  const bool TA = (ta == 'T' || ta == 't' || ta == 'C' || ta == 'c');
  const bool TB = (tb == 'T' || tb == 't' || tb == 'C' || tb == 'c');
  if (k <= 0 || alpha == T(0)) k = 0;  // C <- beta*C (dgemm.f:303-318)
  // C - A*B, the trailing-update form: the update kernels (C read into the
  // accumulators first, A negated)
  if (!TA && !TB && nskip == 0 && k > 0 && alpha == T(-1) && beta == T(1)) {
    launch_update_t<T>(m, n, k, A, lda, B, ldb, C, ldc, 0, s, true);
    return;
  }
  if (!TA && !TB && nskip == 0 && tiles < 64 && k >= 64) {
    const int tm64 = (m + 63) / 64, tn64 = (n + 63) / 64;
    gemm64_kernel<T><<<dim3(tm64 * tn64), 256, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm64);
    return;
  }
  if (!TA && !TB)
    gemm_kernel<T, false, false, false><<<grid, NT, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, chunk, nx);
  else if (!TA && TB)
    gemm_kernel<T, false, true, false><<<grid, NT, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, chunk, nx);
  else if (TA && !TB)
    gemm_kernel<T, true, false, false><<<grid, NT, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, chunk, nx);
  else
    gemm_kernel<T, true, true, false><<<grid, NT, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, chunk, nx);
}

// CUs per XCD a persistent update may hold (LAPACK_AMD_UPD_CUS, default 24
// of 32: at n = 16384, 0.25-0.35 ms per factorization faster than 28; 22 and
// 26 slower than 28, 16 by 5 ms, 32 by 9 ms).
static int update_cus_per_xcd() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("LAPACK_AMD_UPD_CUS");
    v = (e && *e) ? std::atoi(e) : 24;
    if (v < 1) v = 1;
  }
  return v;
}
// The same for FP32 (LAPACK_AMD_UPD_CUS_S), whose update runs twice as fast:
// DSGESV at n = 32768 took 401-404 ms with 20, 404-406 ms with 24, 404 ms
// with 16 and 409 ms with 12.
static int update_cus_per_xcd_s() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("LAPACK_AMD_UPD_CUS_S");
    v = (e && *e) ? std::atoi(e) : 20;
    if (v < 1) v = 1;
  }
  return v;
}

// An update beside the panel below this many flops runs persistent on
// update_cus_per_xcd() CUs per XCD (LAPACK_AMD_UPD_PERSIST_GF, in GFLOP;
// default 60: at n = 16384, 30 is 0.5 ms per factorization slower, 90 and
// 140 are 5 and 13 ms slower).
static double update_persist_flops() {
  static double v = -1;
  if (v < 0) {
    const char* e = std::getenv("LAPACK_AMD_UPD_PERSIST_GF");
    v = 1e9 * ((e && *e) ? std::atof(e) : 60.0);
    if (v < 0) v = 0;
  }
  return v;
}

// Updates with fewer 128x128 tiles than this (and no XCD left to a panel)
// take 64x64 tiles; at n = 16384 any threshold from 128 up gives the same
// factorization time within 0.3 ms.
static int update_small_tiles() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("LAPACK_AMD_UPD64_TILES");
    v = (e && *e) ? std::atoi(e) : 1024;
    if (v < 0) v = 0;
  }
  return v;
}

// Large updates on the software-pipelined kernel (LAPACK_AMD_UPD_PIPE=1) or
// the 256x128 one-tile-per-workgroup kernel (0).
// (read at every launch, so that a process may switch it)
static int update_pipe() {
  const char* e = std::getenv("LAPACK_AMD_UPD_PIPE");
  return (e && *e) ? std::atoi(e) : 0;
}

// Large updates on the direct-to-LDS kernel where its alignment and shape
// conditions hold (LAPACK_AMD_UPD_DMA, default 1; 0 keeps them on
// update_big_kernel).  Measured on MI355X, m = n = 16128, k = 256, alone:
// FP64 47.8 -> 54.8 TFLOP/s, FP32 83.9 -> 119.6 (121.9 at m = n = 32512).
// (read at every launch, so that a process may switch it)
static int update_dma() {
  const char* e = std::getenv("LAPACK_AMD_UPD_DMA");
  return (e && *e) ? std::atoi(e) : 1;
}
// FP32 workgroups of the direct-to-LDS kernel per CU beside a panel
// (LAPACK_AMD_UPD_DMA_OCC, default 1).  With 2, a CU freed by half is
// refilled by the next tile before a panel workgroup (which needs more
// registers than half a CU) fits: SGETRF at n = 32768 took 379 ms with 2
// (panel chain 306 ms) against 321 ms with 1 (panel chain 146 ms).
// k per LDS stage of the FP32 direct-to-LDS kernel beside a panel
// (LAPACK_AMD_UPD_DMA_KS, 16 or 32, default 32; read at every launch).
// Alone 32 is 1-2 % slower (107.3 against 109.3 TFLOP/s at m = n = 16128),
// beside the panel faster: SGETRF 95.2 -> 90.3 ms at n = 16384, 314.2 ->
// 311.7 ms at n = 32768 (the panel's boundary launches wait less).
static int update_dma_ks() {
  const char* e = std::getenv("LAPACK_AMD_UPD_DMA_KS");
  return (e && *e) ? std::atoi(e) : 32;
}
static int update_dma_occ() {
  const char* e = std::getenv("LAPACK_AMD_UPD_DMA_OCC");
  return (e && *e) ? std::atoi(e) : 1;
}
template <class T>
static bool update_dma_fits(int m, int k, const T* A, long lda, const T* B, long ldb, const T* C, long ldc) {
  constexpr long VE = 16 / sizeof(T);
  const unsigned long al = (unsigned long)A | (unsigned long)B | (unsigned long)C;
  return (al & 15) == 0 && m % VE == 0 && k % 16 == 0 && lda % VE == 0 && ldb % VE == 0 && ldc % VE == 0 &&
         (16 * lda + 256) * (long)sizeof(T) < (1L << 31) && (128 * ldb + 16) * (long)sizeof(T) < (1L << 31);
}

// Tiles (256 x 128) from which an update takes the big kernels
// (LAPACK_AMD_UPD_BIG_MIN, default 512; tests lower it to reach those
// kernels at small orders).  Read at every launch.
static long update_big_min() {
  const char* e = std::getenv("LAPACK_AMD_UPD_BIG_MIN");
  return (e && *e) ? std::atol(e) : 512;
}

// alone: no panel runs beside this update (a DGEMM call of the update form)
template <class T>
void launch_update_t(int m, int n, int k, const T* A, long lda, const T* B, long ldb, T* C, long ldc, int nskip,
                     hipStream_t s, bool alone) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  const long big_min = update_big_min();
  const int nx = NXCD - nskip;
  // 256x128 tiles once they still give every CU two workgroups or more.
  // Beside a panel (nskip > 0) an update that is not much longer than the
  // panel chain runs persistent on at most update_cus_per_xcd() CUs of each
  // XCD, so the panel's own short launches find CUs at once instead of
  // queueing behind the update's workgroups.
  const int tmb = (m + 255) / 256, tnb = (n + 127) / 128;
  const long ldmax = std::max(lda, std::max(ldb, ldc)) * (long)sizeof(T);
  if (update_pipe() && (long)tmb * tnb >= big_min && ldmax < (1L << 24)) {
    // the 128 x 128 tiles inside C on the pipelined kernel, the edge strips
    // (fewer than 128 rows or columns) on gemm_kernel
    const int tmp = m / 128, tnp = n / 128;
    const int mi = tmp * 128, ni = tnp * 128;
    const int chunk = (tmp * tnp + nx - 1) / nx;
    int per = std::min(chunk, 32);
    // beside a panel the persistent kernel always leaves CUs of every XCD
    // free: the panel's own launches there would otherwise wait for the
    // whole update
    if (nskip > 0) per = std::min(per, sizeof(T) == 4 ? update_cus_per_xcd_s() : update_cus_per_xcd());
    update_pipe_kernel<T><<<dim3(per * NXCD), 512, 0, s>>>(k, A, lda, B, ldb, C, ldc, tmp, tnp, chunk, nx);
    if (ni < n) {
      const int tm1 = (m + BM - 1) / BM, tn1 = (n - ni + BN - 1) / BN;
      const int ch1 = (tm1 * tn1 + nx - 1) / nx;
      gemm_kernel<T, false, false, true><<<dim3(ch1 * NXCD), NT, 0, s>>>(m, n - ni, k, T(-1), A, lda, B + (long)ni * ldb,
                                                                       ldb, T(1), C + (long)ni * ldc, ldc, tm1, tn1, ch1, nx);
    }
    if (mi < m) {
      const int tm1 = (m - mi + BM - 1) / BM, tn1 = (ni + BN - 1) / BN;
      const int ch1 = (tm1 * tn1 + nx - 1) / nx;
      gemm_kernel<T, false, false, true><<<dim3(ch1 * NXCD), NT, 0, s>>>(m - mi, ni, k, T(-1), A + mi, lda, B, ldb, T(1),
                                                                       C + mi, ldc, tm1, tn1, ch1, nx);
    }
    return;
  }
  if ((long)tmb * tnb >= big_min) {
    const int chunk = (tmb * tnb + nx - 1) / nx;
    int per = chunk;
    if (nskip > 0 && 2.0 * m * (double)n * k < update_persist_flops()) per = std::min(chunk, sizeof(T) == 4 ? update_cus_per_xcd_s() : update_cus_per_xcd());
    if (update_dma() && update_dma_fits<T>(m, k, A, lda, B, ldb, C, ldc)) {
      // FP32: 40 KB of dynamic LDS hold the kernel to one workgroup per CU
      // (two fit otherwise), so a CU freed by a finished tile is free whole
      // for the panel's workgroups instead of taking the next tile at once
      const bool occ1 = sizeof(T) == 4 && !alone && update_dma_occ() == 1;
      if constexpr (sizeof(T) == 4) {
        if (occ1 && k % 32 == 0 && update_dma_ks() == 32) {
          // FP32 beside a panel, stages of 32 k (96 KB of LDS: one workgroup per CU)
          update_dma_kernel<T, 2><<<dim3(per * NXCD), 512, 0, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tmb, tnb, chunk,
                                                                  nx);
          return;
        }
      }
      const size_t dyn = occ1 ? 40 * 1024 : 0;
      update_dma_kernel<T, 1><<<dim3(per * NXCD), 512, dyn, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tmb, tnb, chunk, nx);
      return;
    }
    update_big_kernel<T, 4, 2><<<dim3(per * NXCD), 512, 0, s>>>(m, n, k, A, lda, B, ldb, C, ldc, tmb, tnb, chunk,
                                                               nx);
    return;
  }
  const int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
  if (nskip == 0 && (m <= 64 || tm * tn < update_small_tiles())) {
    // narrow (one leaf's columns), or too small for update_big with the
    // whole GPU to itself (the next panel's columns): 64x64 tiles spread the
    // work over four times the workgroups of 128x128 tiles
    const int tm64 = (m + 63) / 64, tn64 = (n + 63) / 64;
    gemm64_kernel<T, true><<<dim3(tm64 * tn64), 256, 0, s>>>(m, n, k, T(-1), A, lda, B, ldb, T(1), C, ldc, tm64);
    return;
  }
  const int chunk = (tm * tn + nx - 1) / nx;
  gemm_kernel<T, false, false, true><<<dim3(chunk * NXCD), NT, 0, s>>>(m, n, k, T(-1), A, lda, B, ldb, T(1), C, ldc,
                                                                     tm, tn, chunk, nx);
}

}  // namespace

void launch_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, long lda,
                  const double* B, long ldb, double beta, double* C, long ldc, hipStream_t s, int nskip) {
  launch_gemm_t<double>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, nskip);
}
void launch_dgemm_update(int m, int n, int k, const double* A, long lda, const double* B, long ldb, double* C,
                         long ldc, int nskip, hipStream_t s) {
  launch_update_t<double>(m, n, k, A, lda, B, ldb, C, ldc, nskip, s);
}
void launch_sgemm_update(int m, int n, int k, const float* A, long lda, const float* B, long ldb, float* C, long ldc,
                         int nskip, hipStream_t s) {
  launch_update_t<float>(m, n, k, A, lda, B, ldb, C, ldc, nskip, s);
}
void launch_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float* A, long lda,
                  const float* B, long ldb, float beta, float* C, long ldc, hipStream_t s, int nskip) {
  launch_gemm_t<float>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, nskip);
}

}  // namespace la
