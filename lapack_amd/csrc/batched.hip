// batched.hip -- many independent LU factorizations and solves of uneven
// order in one persistent launch: one workgroup per matrix at a time.
//
// The reference has no batched API; its semantics are a loop of DGETRF
// (SRC/dgetrf.f:105-225) and DGETRS (SRC/dgetrs.f:181-218) over the
// matrices, and that is what each workgroup computes for the matrix it takes.
//
// Scheduling: the launch has as many workgroups as fit on the device at once
// (a persistent grid).  Each workgroup takes the next matrix from a work
// counter over an order sorted by decreasing n (largest first, which keeps
// the tail of the launch short), so the grid drains when the counter passes
// the batch size; every loop is bounded by the matrix order.
//
// DGETRF per matrix, right-looking and blocked with panel width NB:
//   1. the panel A(k0:n, k0:k0+w) is loaded into LDS and factored there
//      column by column: IDAMAX (BLAS/SRC/idamax.f:93-122) by wavefront
//      reductions over 64-bit keys that order like its rule (first max,
//      strict '>', a NaN at the head of the column wins, others never), the
//      interchange, the scaling by the reciprocal or by division below sfmin
//      (SRC/dgetrf2.f:196-210) and the rank-1 update of the panel;
//   2. the panel is written back and its interchanges are applied to all
//      other columns (SRC/dgetrf.f:193,199-200): wave 0 first composes the
//      sequence into (position, source) pairs, then each column moves its
//      rows with independent loads followed by stores;
//   3. the columns right of it, BC at a time: the unit lower solve U12 =
//      L11^-1 A12 (dgetrf.f:204-207, the loop order of BLAS/SRC/dtrsm.f:289-300
//      with its zero test), staged in LDS, then the update A22 -= A21*U12
//      (dgetrf.f:212-216) in 4x4 register tiles.
// DGETRS('N') per matrix: the interchanges on B, then blocked forward and
// backward substitution -- one wavefront solves each 64-row diagonal block
// (dtrsm.f:277-300 order, with the reference's zero tests), the workgroup
// updates the rows beyond it.
#include "la_common.h"

namespace la {
namespace {

typedef unsigned long long u64;

template <class T> struct BSfmin;
template <> struct BSfmin<double> { static constexpr double v = 0x1p-1022; };
template <> struct BSfmin<float> { static constexpr float v = 0x1p-126f; };

template <class T>
__device__ __forceinline__ u64 bkey(T x, bool head) {
  const double a = fabs((double)x);
  if (a != a) return head ? 0x7FF0000000000001ull : 1ull;
  return (u64)__double_as_longlong(a) + 1ull;
}

__device__ __forceinline__ void wave_argmax(u64& key, int& row) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const u64 k2 = __shfl_xor(key, off);
    const int r2 = __shfl_xor(row, off);
    if (k2 > key || (k2 == key && r2 < row)) { key = k2; row = r2; }
  }
}

// FP64 / FP32 16x16x4 matrix-core product on gfx950: lane l holds A(l%16,
// l/16) and B(l/16, l%16) of a k-step; accumulator register r of lane l
// holds D(row(l/16, r), l%16).
template <class T> struct BMfma;
template <> struct BMfma<double> {
  typedef la_d4 acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct BMfma<float> {
  typedef la_f4 acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

constexpr int TM = 4, TN = 4, TX = 16;

// Barrier after global-memory writes that other waves of the workgroup read
// next: wait for this wave's stores to complete, then synchronize.
__device__ __forceinline__ void barrier_vm() {
  wait_vm();
  __syncthreads();
}

template <class T, int NB, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NB <= 16 ? 2 : 1))) void getrf_batched_kernel(int batch, const int* __restrict__ nv,
                                                           T* const* __restrict__ av, const int* __restrict__ ldav,
                                                           int* const* __restrict__ ipv, int* __restrict__ infov,
                                                           const int* __restrict__ order, unsigned* counter, int ldp) {
  constexpr int TY = NT / TX;
  constexpr int BC = TY * TN;  // columns per column block
  constexpr int BR = TX * TM;  // rows per row block
  constexpr int NW = NT / 64;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* P = reinterpret_cast<T*>(smem_raw);  // panel, P[c*ldp + r], r relative to k0
  T* U = P + (size_t)NB * ldp;            // U12 column block, U[j*BC + c]
  __shared__ u64 s_key[NW];
  __shared__ int s_row[NW];
  __shared__ int s_piv[NB];
  __shared__ int s_ppos[2 * NB], s_psrc[2 * NB];  // the panel's interchanges as (position, source) pairs
  __shared__ int s_np;
  __shared__ int s_item, s_info;
  const int tid = threadIdx.x;
  const T sfmin = BSfmin<T>::v;

  for (int guard = 0; guard <= batch; ++guard) {
    if (tid == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int item = s_item;
    if (item >= batch) break;
    const int b = order[item];
    const int n = nv[b];
    T* A = av[b];
    const long lda = ldav[b];
    int* ipiv = ipv[b];
    if (tid == 0) s_info = 0;
    for (int k0 = 0; k0 < n; k0 += NB) {
      const int w = min(NB, n - k0);
      const int mr = n - k0;
      // the panel into LDS: a row's NB loads issued together
      for (int r = tid; r < mr; r += NT) {
        T pv[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) pv[c] = c < w ? A[(k0 + r) + (long)(k0 + c) * lda] : T(0);
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < w) P[c * ldp + r] = pv[c];
      }
      __syncthreads();
      // ---- panel in LDS.  Per column: the keys of column j were taken at
      // the end of the previous column's update (or just below for j = 0);
      // B1 -> thread 0 decides -> B2 -> interchange -> B3 -> scale and
      // update (each thread its own rows) fused with the next column's keys.
      u64 key = 0;
      int row = 0x7fffffff;
      for (int r = tid; r < mr; r += NT) {
        const u64 kk = bkey(P[r], r == 0);
        if (kk > key) { key = kk; row = r; }
      }
      for (int j = 0; j < w; ++j) {
        wave_argmax(key, row);
        if ((tid & 63) == 0) { s_key[tid >> 6] = key; s_row[tid >> 6] = row; }
        __syncthreads();
        // every thread folds the waves' candidates (first max, lowest row)
        u64 bk = s_key[0];
        int br = s_row[0];
        for (int q = 1; q < NW; ++q)
          if (s_key[q] > bk || (s_key[q] == bk && s_row[q] < br)) { bk = s_key[q]; br = s_row[q]; }
        if (br >= mr || br < j) br = j;  // cannot happen (row j is always a candidate); keeps indexes in range
        const int p = br;
        // the pivot value and A(j, j), read before anything rewrites column
        // j: the interchange below leaves column j alone, and the update
        // writes it only after that interchange's barrier (or, without an
        // interchange, only below row j = p)
        const T pv = P[j * ldp + p];
        const T ajj = P[j * ldp + j];
        if (tid == 0) {
          s_piv[j] = p;
          ipiv[k0 + j] = k0 + p + 1;
          if (pv == T(0) && s_info == 0) s_info = k0 + j + 1;
        }
        const bool swp = pv != T(0) && p != j;
        if (swp) {
          for (int c = tid; c < w; c += NT) {
            if (c != j) {
              const T t = P[c * ldp + j];
              P[c * ldp + j] = P[c * ldp + p];
              P[c * ldp + p] = t;
            }
          }
          __syncthreads();
          if (tid == 0) P[j * ldp + j] = pv;  // column j's half of the interchange
        }
        const bool recip = fabs(pv) >= sfmin;
        const T rcp = T(1) / pv;
        key = 0;
        row = 0x7fffffff;
        for (int r = j + 1 + tid; r < mr; r += NT) {
          T l = (swp && r == p) ? ajj : P[j * ldp + r];
          if (pv != T(0)) {
            l = recip ? rcp * l : l / pv;
            P[j * ldp + r] = l;
          }
          for (int c = j + 1; c < w; ++c) P[c * ldp + r] = P[c * ldp + r] - l * P[c * ldp + j];
          if (j + 1 < w) {
            const u64 kk = bkey(P[(j + 1) * ldp + r], r == j + 1);
            if (kk > key) { key = kk; row = r; }
          }
        }
        __syncthreads();
      }
      // ---- the panel's interchange sequence as (position, source) pairs:
      // wave 0 replays it over the positions it touches (at most 2w, one
      // per lane), so every column can then move its rows with independent
      // loads followed by stores instead of w dependent swaps.
      if (tid < 64) {
        const int lane = tid;
        int mypos = -1, mysrc = -1, cnt = 0;
        for (int j = 0; j < w; ++j) {
          const int a = k0 + j, bpos = k0 + s_piv[j];
          if (a == bpos) continue;
          const u64 ma = __ballot(mypos == a);
          int ia;
          if (ma == 0ull) {
            if (lane == cnt) { mypos = a; mysrc = a; }
            ia = cnt++;
          } else {
            ia = (int)__builtin_ctzll(ma);
          }
          const u64 mb = __ballot(mypos == bpos);
          int ib;
          if (mb == 0ull) {
            if (lane == cnt) { mypos = bpos; mysrc = bpos; }
            ib = cnt++;
          } else {
            ib = (int)__builtin_ctzll(mb);
          }
          const int sa = __shfl(mysrc, ia), sb = __shfl(mysrc, ib);
          if (lane == ia) mysrc = sb;
          if (lane == ib) mysrc = sa;
        }
        if (lane < cnt) { s_ppos[lane] = mypos; s_psrc[lane] = mysrc; }
        if (lane == 0) s_np = cnt;
      }
      __syncthreads();
      const int np = s_np;
      // interchanges on columns [c_lo, c_hi): a column per thread, its loads
      // issued together, then its stores
      auto swap_cols = [&](int c_lo, int c_hi) {
        for (int c = c_lo + tid; c < c_hi; c += NT) {
          T* col = A + (long)c * lda;
          T tmp[2 * NB];
#pragma unroll
          for (int q = 0; q < 2 * NB; ++q)
            if (q < np) tmp[q] = col[s_psrc[q]];
#pragma unroll
          for (int q = 0; q < 2 * NB; ++q)
            if (q < np) col[s_ppos[q]] = tmp[q];
        }
      };
      // ---- the interchanges on the columns right of the panel, which the
      // solves and the update below read; the panel's write-back and the
      // columns left of it wait until after the update
      if (np > 0) swap_cols(k0 + w, n);
      barrier_vm();  // the solves below read rows other threads moved
      // ---- columns to the right, BC at a time
      const int c1 = k0 + w;
      for (int cb = c1; cb < n; cb += BC) {
        {
          // U12 = L11^-1 A12 for the block's BC = NT/4 columns, four threads
          // per column, thread sub of a column holding rows sub, sub+4, ...;
          // x(j) passes from its owner to the other three by a lane shuffle.
          // Per element the same operations in the same order as the
          // reference's loop (dtrsm.f:289-300, zero test included).
          constexpr int NQ = (NB + 3) / 4;
          const int cl = tid >> 2, sub = tid & 3;
          const int c = cb + cl;
          const bool cok = c < n;
          T* col = A + (long)(cok ? c : 0) * lda;
          T xs[NQ];
#pragma unroll
          for (int m = 0; m < NQ; ++m) {
            const int i = sub + 4 * m;
            xs[m] = (cok && i < w) ? col[k0 + i] : T(0);
          }
#pragma unroll
          for (int j = 0; j < NB; ++j) {
            if (j < w) {
              const T xj = __shfl(xs[j >> 2], (tid & 63 & ~3) | (j & 3));
              if (xj != T(0)) {
#pragma unroll
                for (int m = 0; m < NQ; ++m) {
                  const int i = sub + 4 * m;
                  if (i > j && i < w) xs[m] = xs[m] - xj * P[j * ldp + i];
                }
              }
            }
          }
#pragma unroll
          for (int m = 0; m < NQ; ++m) {
            const int i = sub + 4 * m;
            if (i < NB) {
              if (cok && i < w) col[k0 + i] = xs[m];
              U[i * BC + cl] = (cok && i < w) ? -xs[m] : T(0);  // negated: the update adds
            }
          }
        }
        __syncthreads();  // U complete
        // ---- A22 -= A21*U12 (dgetrf.f:212-216) on the matrix cores, 64-row
        // blocks: wave (wr, wc) owns rows 32 wr .. +31 and columns 32 wc ..
        // +31 of the block (2 x 2 MFMA blocks).  Issued as D^T = U^T A21^T
        // with the accumulator starting from C and U holding -U12, so lane
        // l%16 walks rows and C moves in whole 128-byte pieces; C is loaded
        // before the products.
        {
          typedef BMfma<T> MM;
          typedef typename MM::acc_t acc_t;
          const int lane = tid & 63, wave = tid >> 6;
          const int li = lane & 15, lk = lane >> 4;
          const int wr = wave & 1, wc = wave >> 1;
          const int nks = (w + 3) >> 2;
          for (int rb = c1; rb < n; rb += 64) {
            const int row0 = rb + 32 * wr;
            acc_t acc[2][2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) {
              const int row = row0 + 16 * bi + li;
#pragma unroll
              for (int bj = 0; bj < 2; ++bj) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  const int col = cb + 32 * wc + 16 * bj + MM::row(lk, r);
                  acc[bi][bj][r] = (row < n && col < n) ? A[row + (long)col * lda] : T(0);
                }
              }
            }
            for (int s4 = 0; s4 < nks; ++s4) {
              const int k = 4 * s4 + lk;
              T af[2], bf[2];
#pragma unroll
              for (int bj = 0; bj < 2; ++bj) af[bj] = k < w ? U[k * BC + 32 * wc + 16 * bj + li] : T(0);
#pragma unroll
              for (int bi = 0; bi < 2; ++bi) {
                const int row = row0 + 16 * bi + li;
                bf[bi] = (k < w && row < n) ? P[k * ldp + (row - k0)] : T(0);
              }
#pragma unroll
              for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = MM::op(af[bj], bf[bi], acc[bi][bj]);
            }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) {
              const int row = row0 + 16 * bi + li;
#pragma unroll
              for (int bj = 0; bj < 2; ++bj) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  const int col = cb + 32 * wc + 16 * bj + MM::row(lk, r);
                  if (row < n && col < n) A[row + (long)col * lda] = acc[bi][bj][r];
                }
              }
            }
          }
        }
        __syncthreads();
      }
      // ---- panel back to HBM; the interchanges on the columns left of it
      for (int r = tid; r < mr; r += NT)
        for (int c = 0; c < w; ++c) A[(k0 + r) + (long)(k0 + c) * lda] = P[c * ldp + r];
      if (np > 0) swap_cols(0, k0);
      barrier_vm();  // the next panel loads what this one's update wrote
    }
    if (tid == 0) infov[b] = s_info;
    __syncthreads();
  }
}

// DGETRS('N') for one matrix per workgroup, NR right-hand sides at a time.
template <class T, int NT, int NR>
__global__ __launch_bounds__(NT) void getrs_batched_kernel(int batch, const int* __restrict__ nv,
                                                           const int* __restrict__ nrhsv, T* const* __restrict__ av,
                                                           const int* __restrict__ ldav, int* const* __restrict__ ipv,
                                                           T* const* __restrict__ bv, const int* __restrict__ ldbv,
                                                           const int* __restrict__ order, unsigned* counter, int ldx) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* X = reinterpret_cast<T*>(smem_raw);  // X[r + q*ldx]
  __shared__ int s_item;
  const int tid = threadIdx.x, lane = tid & 63;
  for (int guard = 0; guard <= batch; ++guard) {
    if (tid == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int item = s_item;
    if (item >= batch) break;
    const int b = order[item];
    const int n = nv[b], nrhs = nrhsv[b];
    const T* A = av[b];
    const long lda = ldav[b];
    const int* ipiv = ipv[b];
    T* B = bv[b];
    const long ldb = ldbv[b];
    for (int q0 = 0; q0 < nrhs; q0 += NR) {
      const int nr = min(NR, nrhs - q0);
      for (int q = 0; q < nr; ++q)
        for (int r = tid; r < n; r += NT) X[r + q * ldx] = B[r + (long)(q0 + q) * ldb];
      __syncthreads();
      if (tid < nr) {  // interchanges, in order (SRC/dgetrs.f:187)
        for (int k = 0; k < n; ++k) {
          int p = ipiv[k] - 1;
          if (p < 0 || p >= n) p = k;
          if (p != k) { const T t = X[k + tid * ldx]; X[k + tid * ldx] = X[p + tid * ldx]; X[p + tid * ldx] = t; }
        }
      }
      __syncthreads();
      // L Y = X, unit lower (dgetrs.f:191)
      for (int kb = 0; kb < n; kb += 64) {
        const int bs = min(64, n - kb);
        if (tid < 64) {
          T xi[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) xi[q] = (q < nr && lane < bs) ? X[kb + lane + q * ldx] : T(0);
          for (int k = 0; k < bs; ++k) {
            const T l = (lane > k && lane < bs) ? A[(kb + lane) + (long)(kb + k) * lda] : T(0);
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              const T xk = __shfl(xi[q], k);
              if (xk != T(0) && lane > k && lane < bs) xi[q] = xi[q] - xk * l;
            }
          }
#pragma unroll
          for (int q = 0; q < NR; ++q)
            if (q < nr && lane < bs) X[kb + lane + q * ldx] = xi[q];
        }
        __syncthreads();
        for (int i = kb + bs + tid; i < n; i += NT) {
          T xi[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) xi[q] = (q < nr) ? X[i + q * ldx] : T(0);
          for (int k = 0; k < bs; ++k) {
            const T l = A[i + (long)(kb + k) * lda];
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              const T xk = (q < nr) ? X[kb + k + q * ldx] : T(0);
              if (xk != T(0)) xi[q] = xi[q] - xk * l;
            }
          }
#pragma unroll
          for (int q = 0; q < NR; ++q)
            if (q < nr) X[i + q * ldx] = xi[q];
        }
        __syncthreads();
      }
      // U X = Y, non-unit upper (dgetrs.f:197)
      for (int kb = ((n - 1) / 64) * 64; kb >= 0; kb -= 64) {
        const int bs = min(64, n - kb);
        if (tid < 64) {
          T xi[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) xi[q] = (q < nr && lane < bs) ? X[kb + lane + q * ldx] : T(0);
          const T dii = (lane < bs) ? A[(kb + lane) + (long)(kb + lane) * lda] : T(1);
          for (int k = bs - 1; k >= 0; --k) {
            const T u = (lane < k) ? A[(kb + lane) + (long)(kb + k) * lda] : T(0);
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              if (lane == k && xi[q] != T(0)) xi[q] = xi[q] / dii;
              const T xk = __shfl(xi[q], k);
              if (xk != T(0) && lane < k) xi[q] = xi[q] - xk * u;
            }
          }
#pragma unroll
          for (int q = 0; q < NR; ++q)
            if (q < nr && lane < bs) X[kb + lane + q * ldx] = xi[q];
        }
        __syncthreads();
        for (int i = tid; i < kb; i += NT) {
          T xi[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) xi[q] = (q < nr) ? X[i + q * ldx] : T(0);
          for (int k = bs - 1; k >= 0; --k) {
            const T u = A[i + (long)(kb + k) * lda];
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              const T xk = (q < nr) ? X[kb + k + q * ldx] : T(0);
              if (xk != T(0)) xi[q] = xi[q] - xk * u;
            }
          }
#pragma unroll
          for (int q = 0; q < NR; ++q)
            if (q < nr) X[i + q * ldx] = xi[q];
        }
        __syncthreads();
      }
      for (int q = 0; q < nr; ++q)
        for (int r = tid; r < n; r += NT) B[r + (long)(q0 + q) * ldb] = X[r + q * ldx];
      __syncthreads();
    }
  }
}

// DGETRS('T') for one matrix per workgroup, NR right-hand sides at a time
// (SRC/dgetrs.f:199-217): U^T Y = B (forward), L^T X = Y (backward, unit
// diagonal), then the interchanges in reverse order.  The right-hand sides
// live in LDS.  Each 64-row block first takes the contributions of the rows
// already solved as dot products down columns of A (row i of U^T is column
// i of A, read coalesced, one wavefront per column, a fixed-order wavefront
// reduction), then one wavefront solves the block's 64 x 64 triangle from a
// copy staged in LDS.
template <class T, int NT, int NR>
__global__ __launch_bounds__(NT) void getrs_t_batched_kernel(int batch, const int* __restrict__ nv,
                                                             const int* __restrict__ nrhsv, T* const* __restrict__ av,
                                                             const int* __restrict__ ldav, int* const* __restrict__ ipv,
                                                             T* const* __restrict__ bv, const int* __restrict__ ldbv,
                                                             const int* __restrict__ order, unsigned* counter,
                                                             int ldx) {
  constexpr int NW = NT / 64;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* X = reinterpret_cast<T*>(smem_raw);  // X[r + q*ldx]
  __shared__ T sD[64][65];                // sD[c][r] = A(kb + r, kb + c)
  __shared__ int s_item;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int guard = 0; guard <= batch; ++guard) {
    if (tid == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int item = s_item;
    if (item >= batch) break;
    const int b = order[item];
    const int n = nv[b], nrhs = nrhsv[b];
    const T* A = av[b];
    const long lda = ldav[b];
    const int* ipiv = ipv[b];
    T* B = bv[b];
    const long ldb = ldbv[b];
    auto stage_diag = [&](int kb, int bs) {
      for (int e = tid; e < 64 * 64; e += NT) {
        const int r = e & 63, c = e >> 6;
        sD[c][r] = (r < bs && c < bs) ? A[(kb + r) + (long)(kb + c) * lda] : T(0);
      }
    };
    for (int q0 = 0; q0 < nrhs; q0 += NR) {
      const int nr = min(NR, nrhs - q0);
      for (int q = 0; q < nr; ++q)
        for (int r = tid; r < n; r += NT) X[r + q * ldx] = B[r + (long)(q0 + q) * ldb];
      __syncthreads();
      // U^T Y = B (dgetrs.f:205)
      for (int kb = 0; kb < n; kb += 64) {
        const int bs = min(64, n - kb);
        for (int c = wave; c < bs; c += NW) {
          const int i = kb + c;
          T acc[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = T(0);
          for (int k = lane; k < kb; k += 64) {
            const T u = A[k + (long)i * lda];
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] = acc[q] + u * ((q < nr) ? X[k + q * ldx] : T(0));
          }
#pragma unroll
          for (int q = 0; q < NR; ++q)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[q] = acc[q] + __shfl_xor(acc[q], off);
          if (lane == 0)
            for (int q = 0; q < nr; ++q) X[i + q * ldx] = X[i + q * ldx] - acc[q];
        }
        stage_diag(kb, bs);
        __syncthreads();
        if (wave == 0) {
          T xi[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) xi[q] = (q < nr && lane < bs) ? X[kb + lane + q * ldx] : T(0);
          for (int k = 0; k < bs; ++k) {
            const T dkk = sD[k][k];
            const T u = (lane > k && lane < bs) ? sD[lane][k] : T(0);  // U(kb+k, kb+lane)
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              if (lane == k) xi[q] = xi[q] / dkk;
              const T xk = __shfl(xi[q], k);
              if (lane > k && lane < bs) xi[q] = xi[q] - u * xk;
            }
          }
#pragma unroll
          for (int q = 0; q < NR; ++q)
            if (q < nr && lane < bs) X[kb + lane + q * ldx] = xi[q];
        }
        __syncthreads();
      }
      // L^T X = Y, unit diagonal (dgetrs.f:211)
      for (int kb = ((n - 1) / 64) * 64; kb >= 0; kb -= 64) {
        const int bs = min(64, n - kb);
        for (int c = wave; c < bs; c += NW) {
          const int i = kb + c;
          T acc[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = T(0);
          for (int k = kb + bs + lane; k < n; k += 64) {
            const T l = A[k + (long)i * lda];
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] = acc[q] + l * ((q < nr) ? X[k + q * ldx] : T(0));
          }
#pragma unroll
          for (int q = 0; q < NR; ++q)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[q] = acc[q] + __shfl_xor(acc[q], off);
          if (lane == 0)
            for (int q = 0; q < nr; ++q) X[i + q * ldx] = X[i + q * ldx] - acc[q];
        }
        stage_diag(kb, bs);
        __syncthreads();
        if (wave == 0) {
          T xi[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) xi[q] = (q < nr && lane < bs) ? X[kb + lane + q * ldx] : T(0);
          for (int k = bs - 1; k >= 0; --k) {
            const T l = (lane < k) ? sD[lane][k] : T(0);  // L(kb+k, kb+lane)
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              const T xk = __shfl(xi[q], k);
              if (lane < k) xi[q] = xi[q] - l * xk;
            }
          }
#pragma unroll
          for (int q = 0; q < NR; ++q)
            if (q < nr && lane < bs) X[kb + lane + q * ldx] = xi[q];
        }
        __syncthreads();
      }
      if (tid < nr) {  // interchanges in reverse order (dgetrs.f:217)
        for (int k = n - 1; k >= 0; --k) {
          int p = ipiv[k] - 1;
          if (p < 0 || p >= n) p = k;
          if (p != k) { const T t = X[k + tid * ldx]; X[k + tid * ldx] = X[p + tid * ldx]; X[p + tid * ldx] = t; }
        }
      }
      __syncthreads();
      for (int q = 0; q < nr; ++q)
        for (int r = tid; r < n; r += NT) B[r + (long)(q0 + q) * ldb] = X[r + q * ldx];
      __syncthreads();
    }
  }
}

// Abar_b <- -Bbar_b X_b^T (+ Abar_b when accumulate), for every matrix b of
// the batch: the adjoint's rank-nrhs product, bound by the n^2 result
// written to HBM.  Grid (tiles of the largest order, batch): a 64 x 64 tile
// per 256-thread workgroup, 32 of the nrhs columns at a time staged in LDS,
// each lane writing 16 columns of one row; tiles past a matrix's order exit.
__global__ __launch_bounds__(256) void rankk_batched_kernel(const int* __restrict__ nv, const int* __restrict__ nrhsv,
                                                            double* const* __restrict__ bbv,
                                                            const int* __restrict__ ldbbv,
                                                            double* const* __restrict__ xv, const int* __restrict__ ldxv,
                                                            double* const* __restrict__ cv,
                                                            const int* __restrict__ ldcv, int tiles1, int accumulate) {
  constexpr int TL = 64, KC = 32;
  __shared__ double sa[KC][TL];
  __shared__ double sb[KC][TL + 1];
  const int b = blockIdx.y;
  const int n = nv[b];
  const int i0 = (blockIdx.x % tiles1) * TL, j0 = (blockIdx.x / tiles1) * TL;
  if (i0 >= n || j0 >= n) return;
  const int k = nrhsv[b];
  const double* A = bbv[b];
  const long lda = ldbbv[b];
  const double* Bm = xv[b];
  const long ldbm = ldxv[b];
  double* C = cv[b];
  const long ldc = ldcv[b];
  const int t = threadIdx.x;
  const int r = t % TL, cg = t / TL;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  for (int p0 = 0; p0 < k; p0 += KC) {
    const int kc = min(KC, k - p0);
    __syncthreads();
    for (int e = t; e < kc * TL; e += 256) {
      const int p = e / TL, rr = e % TL;
      sa[p][rr] = (i0 + rr < n) ? A[(i0 + rr) + (long)(p0 + p) * lda] : 0.0;
      sb[p][rr] = (j0 + rr < n) ? Bm[(j0 + rr) + (long)(p0 + p) * ldbm] : 0.0;
    }
    __syncthreads();
    for (int p = 0; p < kc; ++p) {
      const double a = sa[p][r];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] = fma(a, sb[p][cg * 16 + q], acc[q]);
    }
  }
  const int i = i0 + r;
  if (i >= n) return;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int j = j0 + cg * 16 + q;
    if (j < n) {
      double* cp = C + i + (long)j * ldc;
      *cp = accumulate ? *cp - acc[q] : -acc[q];
    }
  }
}

// uniform(-1,1) from a hash of the element index: synthetic input of the
// autotuner's timing runs.
template <class T>
__global__ void fill_uniform_kernel(T* a, size_t count, unsigned long long seed) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    unsigned long long z = seed + i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    a[i] = (T)((double)(z >> 11) * 0x1p-52 - 1.0);
  }
}

template <class T, int NB, int NT>
size_t getrf_smem(int ldp) {
  return sizeof(T) * ((size_t)NB * ldp + (size_t)NB * (NT / TX) * TN);
}

// Workgroups of the persistent grid: all that fit on the device at once.
template <class T, int NB, int NT>
int grid_cfg(int batch, int nmax) {
  const int ldp = (nmax + 3) & ~3;
  const size_t sm = getrf_smem<T, NB, NT>(ldp);
  if (sm + 1024 > (size_t)lds_limit()) return LA_ERR_LIMIT;  // static LDS of the kernel included
  auto kern = getrf_batched_kernel<T, NB, NT>;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit()) != hipSuccess)
    (void)hipGetLastError();
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, sm) != hipSuccess || occ < 1) {
    (void)hipGetLastError();
    occ = 1;
  }
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  return std::max(1, std::min(batch, ncu * occ));
}

template <class T, int NB, int NT>
void launch_getrf_cfg(int grid, int batch, int nmax, const int* nv, T* const* av, const int* ldav, int* const* ipv,
                      int* infov, const int* order, unsigned* counter, hipStream_t s) {
  const int ldp = (nmax + 3) & ~3;  // 32-byte aligned columns for the tile loads
  const size_t sm = getrf_smem<T, NB, NT>(ldp);
  launch_fill_u32(counter, 1, 0u, s);
  getrf_batched_kernel<T, NB, NT><<<grid, NT, sm, s>>>(batch, nv, av, ldav, ipv, infov, order, counter, ldp);
}

}  // namespace

void launch_fill_uniform_d(double* a, size_t count, unsigned long long seed, hipStream_t s) {
  if (count) fill_uniform_kernel<double><<<1024, 256, 0, s>>>(a, count, seed);
}
void launch_fill_uniform_s(float* a, size_t count, unsigned long long seed, hipStream_t s) {
  if (count) fill_uniform_kernel<float><<<1024, 256, 0, s>>>(a, count, seed);
}

// Configurations the autotuner chooses from: (NB, threads).
int batched_getrf_configs() { return 6; }
void batched_getrf_config(int idx, int* nb, int* nt) {
  static const int tab[6][2] = {{8, 256}, {16, 256}, {16, 512}, {32, 256}, {16, 128}, {8, 128}};
  *nb = tab[idx][0];
  *nt = tab[idx][1];
}
size_t batched_getrf_smem(int idx, int nmax, int elem) {
  int nb, nt;
  batched_getrf_config(idx, &nb, &nt);
  // + 1 KiB for the kernel's static LDS (pivot keys, rows, indexes)
  return (size_t)elem * ((size_t)nb * ((nmax + 3) & ~3) + (size_t)nb * (nt / TX) * TN) + 1024;
}

template <class T>
int batched_getrf_grid(int cfg, int batch, int nmax) {
  switch (cfg) {
    case 0: return grid_cfg<T, 8, 256>(batch, nmax);
    case 1: return grid_cfg<T, 16, 256>(batch, nmax);
    case 2: return grid_cfg<T, 16, 512>(batch, nmax);
    case 3: return grid_cfg<T, 32, 256>(batch, nmax);
    case 4: return grid_cfg<T, 16, 128>(batch, nmax);
    case 5: return grid_cfg<T, 8, 128>(batch, nmax);
  }
  return LA_ERR_LIMIT;
}
template int batched_getrf_grid<double>(int, int, int);
template int batched_getrf_grid<float>(int, int, int);

template <class T>
int launch_getrf_batched(int cfg, int grid, int batch, int nmax, const int* nv, T* const* av, const int* ldav,
                         int* const* ipv, int* infov, const int* order, unsigned* counter, hipStream_t s) {
  switch (cfg) {
    case 0: launch_getrf_cfg<T, 8, 256>(grid, batch, nmax, nv, av, ldav, ipv, infov, order, counter, s); return 0;
    case 1: launch_getrf_cfg<T, 16, 256>(grid, batch, nmax, nv, av, ldav, ipv, infov, order, counter, s); return 0;
    case 2: launch_getrf_cfg<T, 16, 512>(grid, batch, nmax, nv, av, ldav, ipv, infov, order, counter, s); return 0;
    case 3: launch_getrf_cfg<T, 32, 256>(grid, batch, nmax, nv, av, ldav, ipv, infov, order, counter, s); return 0;
    case 4: launch_getrf_cfg<T, 16, 128>(grid, batch, nmax, nv, av, ldav, ipv, infov, order, counter, s); return 0;
    case 5: launch_getrf_cfg<T, 8, 128>(grid, batch, nmax, nv, av, ldav, ipv, infov, order, counter, s); return 0;
  }
  return LA_ERR_LIMIT;
}
template int launch_getrf_batched<double>(int, int, int, int, const int*, double* const*, const int*, int* const*,
                                          int*, const int*, unsigned*, hipStream_t);
template int launch_getrf_batched<float>(int, int, int, int, const int*, float* const*, const int*, int* const*,
                                         int*, const int*, unsigned*, hipStream_t);

template <class T>
int batched_getrs_grid(int batch, int nmax) {
  constexpr int NT = 256, NR = 4;
  const size_t sm = sizeof(T) * (size_t)NR * std::max(nmax, 1);
  if (sm > (size_t)lds_limit()) return LA_ERR_LIMIT;
  auto kern = getrs_batched_kernel<T, NT, NR>;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit()) != hipSuccess)
    (void)hipGetLastError();
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, sm) != hipSuccess || occ < 1) {
    (void)hipGetLastError();
    occ = 1;
  }
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  return std::max(1, std::min(batch, ncu * occ));
}
template int batched_getrs_grid<double>(int, int);
template int batched_getrs_grid<float>(int, int);

template <class T>
int launch_getrs_batched(int grid, int batch, int nmax, const int* nv, const int* nrhsv, T* const* av,
                         const int* ldav, int* const* ipv, T* const* bv, const int* ldbv, const int* order,
                         unsigned* counter, hipStream_t s) {
  constexpr int NT = 256, NR = 4;
  const size_t sm = sizeof(T) * (size_t)NR * std::max(nmax, 1);
  launch_fill_u32(counter, 1, 0u, s);
  getrs_batched_kernel<T, NT, NR><<<grid, NT, sm, s>>>(batch, nv, nrhsv, av, ldav, ipv, bv, ldbv, order, counter,
                                                      std::max(nmax, 1));
  return 0;
}
template <class T>
int batched_getrs_t_grid(int batch, int nmax) {
  constexpr int NT = 256, NR = 4;
  const size_t sm = sizeof(T) * (size_t)NR * std::max(nmax, 1);
  if (sm + sizeof(T) * 64 * 65 + 64 > (size_t)lds_limit()) return LA_ERR_LIMIT;
  auto kern = getrs_t_batched_kernel<T, NT, NR>;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit()) != hipSuccess)
    (void)hipGetLastError();
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, sm) != hipSuccess || occ < 1) {
    (void)hipGetLastError();
    occ = 1;
  }
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  return std::max(1, std::min(batch, ncu * occ));
}
template int batched_getrs_t_grid<double>(int, int);

// The batched adjoint: Bbar = A^-T Xbar in place of Xbar (getrs_t_batched),
// then Abar = -Bbar X^T (+ Abar).
int launch_adjoint_batched(int grid, int batch, int nmax, const int* nv, const int* nrhsv, double* const* luv,
                           const int* ldluv, int* const* ipv, double* const* xv, const int* ldxv,
                           double* const* xbv, const int* ldxbv, double* const* abv, const int* ldabv,
                           int accumulate, const int* order, unsigned* counter, hipStream_t s) {
  constexpr int NT = 256, NR = 4;
  const size_t sm = sizeof(double) * (size_t)NR * std::max(nmax, 1);
  launch_fill_u32(counter, 1, 0u, s);
  getrs_t_batched_kernel<double, NT, NR><<<grid, NT, sm, s>>>(batch, nv, nrhsv, luv, ldluv, ipv, xbv, ldxbv, order,
                                                             counter, std::max(nmax, 1));
  const int tiles1 = (nmax + 63) / 64;
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    const int nb = std::min(65535, batch - b0);
    rankk_batched_kernel<<<dim3(tiles1 * tiles1, nb), 256, 0, s>>>(nv + b0, nrhsv + b0, xbv + b0, ldxbv + b0, xv + b0,
                                                                 ldxv + b0, abv + b0, ldabv + b0, tiles1, accumulate);
  }
  return 0;
}

template int launch_getrs_batched<double>(int, int, int, const int*, const int*, double* const*, const int*,
                                          int* const*, double* const*, const int*, const int*, unsigned*, hipStream_t);
template int launch_getrs_batched<float>(int, int, int, const int*, const int*, float* const*, const int*,
                                         int* const*, float* const*, const int*, const int*, unsigned*, hipStream_t);

}  // namespace la
