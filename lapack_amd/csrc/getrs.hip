// getrs.hip -- the triangular solves of DGETRS as one launch per sweep.
//
// DGETRS (SRC/dgetrs.f:181-218) solves with L (unit lower) then U (upper),
// or U^T then L^T for 'T'.  Each sweep is M x = y for one triangular M =
// op(tri(A)) over the n x nrhs right-hand sides X (nrhs <= 16 per launch).
// A sweep is bound by reading the triangle once (4 n^2 bytes in FP64) and
// by the chain through the diagonal blocks, so it gets its own kernel:
//
//   * one 256-thread workgroup per 64-row block of x, ordered by a ticket
//     taken at start: a workgroup only waits for blocks of lower tickets,
//     i.e. for workgroups already running, so every wait ends;
//   * thread (i, g) owns row i of the block and the columns c in
//     [16g, 16g+16) of every 64-column tile of M it depends on; tiles are
//     read straight from HBM, four dependency blocks per round: their
//     flags are polled together, their x blocks staged in LDS together, and
//     the thread accumulates acc(q) -= sum_c M(i,c) x(c,q) for all q;
//   * the four partial sums per row are added in LDS, y = b + acc, and the
//     diagonal block is solved by substitution, one wavefront per four
//     right-hand sides, lane i holding row i: for each k, x_k is finished
//     (divided by the diagonal when M is not unit) and broadcast, and the
//     rows below (forward) or above (backward) subtract M(i,k) x_k -- the
//     column-oriented order of BLAS/SRC/dtrsm.f:277-300, with its zero test;
//   * x is stored with device-coherent stores, the stores are waited for,
//     and the block's flag is raised.
#include "la_common.h"

namespace la {
namespace {

constexpr int TB = 64;    // rows per block
constexpr int TQ = 16;    // right-hand sides per launch
constexpr int DEPB = 4;   // dependency blocks per round

template <bool TRANS>
__device__ __forceinline__ long m_index(int i, int c, long lda) {
  return TRANS ? (long)c + (long)i * lda : (long)i + (long)c * lda;
}

template <class T, bool LOWERM, bool UNIT, bool TRANS>
__global__ __launch_bounds__(256) void tri_chain_kernel(int n, int nrhs, const T* __restrict__ A, long lda,
                                                        T* __restrict__ X, long ldx, unsigned* ticket,
                                                        unsigned* flags, int* err) {
  __shared__ T sD[TB][TB + 1];             // diagonal block of M
  __shared__ T sx[DEPB][TB][TQ];           // x blocks of a round; reused for the partial sums
  __shared__ int s_t;
  const int tid = threadIdx.x;
  const int nblk = (n + TB - 1) / TB;
  if (tid == 0) s_t = (int)__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int t = s_t;
  if (t >= nblk) return;
  const int bi = LOWERM ? t : nblk - 1 - t;
  const int r0 = bi * TB;
  const int bs = min(TB, n - r0);

  for (int e = tid; e < TB * TB; e += 256) {
    int i, c;
    if (TRANS) { c = e % TB; i = e / TB; } else { i = e % TB; c = e / TB; }
    sD[i][c] = (i < bs && c < bs) ? A[m_index<TRANS>(r0 + i, r0 + c, lda)] : T(0);
  }

  const int i = tid & (TB - 1);
  const int g = tid >> 6;
  T acc[TQ];
#pragma unroll
  for (int q = 0; q < TQ; ++q) acc[q] = T(0);
  const int ndep = LOWERM ? bi : nblk - 1 - bi;
  bool failed = false;
  for (int d0 = 0; d0 < ndep; d0 += DEPB) {
    const int nd = min(DEPB, ndep - d0);
    // M values of this round, issued before the waits
    T mv[DEPB][16];
#pragma unroll
    for (int u = 0; u < DEPB; ++u) {
      const int j = LOWERM ? d0 + u : nblk - 1 - (d0 + u);
      const int c0 = j * TB + 16 * g;
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        T v = T(0);
        if (u < nd && i < bs && c0 + c < n) v = A[m_index<TRANS>(r0 + i, c0 + c, lda)];
        mv[u][c] = v;
      }
    }
    if (tid < nd && !failed) {
      const int j = LOWERM ? d0 + tid : nblk - 1 - (d0 + tid);
      unsigned spins = 0;
      while (ld_dev(&flags[j]) == 0u) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > (1u << 22)) { st_dev(err, 1); failed = true; break; }
      }
    }
    __syncthreads();
    for (int e = tid; e < DEPB * TB * TQ; e += 256) {
      const int c = e % TB, q = (e / TB) % TQ, u = e / (TB * TQ);
      const int j = LOWERM ? d0 + u : nblk - 1 - (d0 + u);
      T v = T(0);
      if (u < nd && q < nrhs && j * TB + c < n) v = ld_dev(&X[(long)(j * TB + c) + (long)q * ldx]);
      sx[u][c][q] = v;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < DEPB; ++u) {
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const T m = mv[u][c];
#pragma unroll
        for (int q = 0; q < TQ; ++q) acc[q] = acc[q] - m * sx[u][16 * g + c][q];
      }
    }
    __syncthreads();
  }

  // partial sums of the four column groups, then y = b + acc
  T* red = &sx[0][0][0];  // [4][TB][TQ]
#pragma unroll
  for (int q = 0; q < TQ; ++q) red[(g * TB + i) * TQ + q] = acc[q];
  __syncthreads();
  // wave g solves right-hand sides 4g .. 4g+3 of the diagonal block
  const int lane = i;
  T x[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int q = 4 * g + u;
    T y = T(0);
    if (lane < bs && q < nrhs) {
      y = X[(long)(r0 + lane) + (long)q * ldx];
      y = y + (((red[(0 * TB + lane) * TQ + q] + red[(1 * TB + lane) * TQ + q]) + red[(2 * TB + lane) * TQ + q]) +
               red[(3 * TB + lane) * TQ + q]);
    }
    x[u] = y;
  }
  if (LOWERM) {
    for (int k = 0; k < bs; ++k) {
      const T mik = sD[lane][k];
      const T dkk = sD[k][k];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!UNIT && lane == k && x[u] != T(0)) x[u] = x[u] / dkk;
        const T xk = __shfl(x[u], k);
        if (xk != T(0) && lane > k) x[u] = x[u] - xk * mik;
      }
    }
  } else {
    for (int k = bs - 1; k >= 0; --k) {
      const T mik = sD[lane][k];
      const T dkk = sD[k][k];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!UNIT && lane == k && x[u] != T(0)) x[u] = x[u] / dkk;
        const T xk = __shfl(x[u], k);
        if (xk != T(0) && lane < k) x[u] = x[u] - xk * mik;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int q = 4 * g + u;
    if (lane < bs && q < nrhs) st_dev(&X[(long)(r0 + lane) + (long)q * ldx], x[u]);
  }
  wait_vm();
  __syncthreads();
  if (tid == 0) st_dev(&flags[bi], 1u);
}

template <class T, bool LOWERM, bool UNIT, bool TRANS>
void launch_chain(int n, int nrhs, const T* A, long lda, T* X, long ldx, unsigned* ctl, int* err, hipStream_t s) {
  const int nblk = (n + TB - 1) / TB;
  for (int q0 = 0; q0 < nrhs; q0 += TQ) {
    const int nq = min(TQ, nrhs - q0);
    launch_fill_u32(ctl, (size_t)nblk + 1, 0u, s);  // ticket + flags (a kernel: graph-safe)
    tri_chain_kernel<T, LOWERM, UNIT, TRANS><<<nblk, 256, 0, s>>>(n, nq, A, lda, X + (long)q0 * ldx, ldx, ctl,
                                                                  ctl + 1, err);
  }
}

}  // namespace

size_t tri_chain_ctl_words(int n) { return (size_t)(n + TB - 1) / TB + 1; }

// which: 0 = L (unit lower), 1 = U (upper), 2 = U^T, 3 = L^T.
template <class T>
void launch_tri_chain(int which, int n, int nrhs, const T* A, long lda, T* X, long ldx, unsigned* ctl, int* err,
                      hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  switch (which) {
    case 0: launch_chain<T, true, true, false>(n, nrhs, A, lda, X, ldx, ctl, err, s); break;
    case 1: launch_chain<T, false, false, false>(n, nrhs, A, lda, X, ldx, ctl, err, s); break;
    case 2: launch_chain<T, true, false, true>(n, nrhs, A, lda, X, ldx, ctl, err, s); break;
    case 3: launch_chain<T, false, true, true>(n, nrhs, A, lda, X, ldx, ctl, err, s); break;
  }
}
template void launch_tri_chain<double>(int, int, int, const double*, long, double*, long, unsigned*, int*,
                                       hipStream_t);
template void launch_tri_chain<float>(int, int, int, const float*, long, float*, long, unsigned*, int*,
                                      hipStream_t);

}  // namespace la
