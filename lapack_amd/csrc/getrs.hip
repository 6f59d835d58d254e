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
//     read straight from HBM, up to two dependency blocks per round: their
//     flags are polled together, their x blocks staged in LDS together, and
//     the round's products are summed on their own before they join the
//     running sum (two-level summation: no sum runs over more than 64 terms
//     of one round or one term per round).  The block's last dependency --
//     the one finished most recently -- gets a round of its own, so the
//     work left once it arrives is one tile;
//   * before its dependencies arrive, the workgroup inverts the 64x64
//     diagonal block in LDS (doubling over 2x2 block-triangular inverses),
//     so that once they have arrived the block costs a 64-term product per
//     row, not a 64-step chain;
//   * the four partial sums per row are added in LDS, y = b + acc, and
//     x = inv(M_dd) y, one wavefront per four right-hand sides;
//   * x is stored with device-coherent stores, the stores are waited for,
//     and the block's flag is raised.
// Every wait is bounded; one that gives up sets *err, and the other waits
// poll *err and give up as well, so every workgroup runs to its end (the
// caller then leaves B as it was and reports the failure).
#include "la_common.h"

namespace la {
namespace {

constexpr int TB = 64;    // rows per block
constexpr int TQ = 16;    // right-hand sides per launch (at most)
constexpr int DEPB = 2;   // dependency blocks per round

template <bool TRANS>
__device__ __forceinline__ long m_index(int i, int c, long lda) {
  return TRANS ? (long)c + (long)i * lda : (long)i + (long)c * lda;
}

// Replaces the TB x TB triangle in sD (NTH threads; unit diagonal when
// UNIT) by its inverse, by doubling: the 1x1 blocks are the reciprocals of
// the diagonal (1 when unit); then for b = 1, 2, .., 32 every pair of
// inverted b x b diagonal blocks A (first) and B (second) gets its
// off-diagonal block, -inv(B) C inv(A) below the diagonal (lower M) or
// -inv(A) C inv(B) above it (upper M), through a b x b temporary tmp
// (>= TB*TB/4 values).  Ends with a barrier.
template <class T, bool LOWERM, bool UNIT, int NTH = 256>
__device__ void invert_diag(T (*sD)[TB + 1], T* tmp) {
  const int tid = threadIdx.x;
  if (tid < TB) sD[tid][tid] = UNIT ? T(1) : T(1) / sD[tid][tid];
  __syncthreads();
  for (int bsz = 1; bsz < TB; bsz *= 2) {
    const int cnt = (TB / (2 * bsz)) * bsz * bsz;  // elements of all pairs' blocks
    // phase 1: tmp = C inv(A) (lower) or inv(A) C (upper), per pair
    for (int e = tid; e < cnt; e += NTH) {
      const int pr = e / (bsz * bsz), r = (e / bsz) % bsz, c = e % bsz;
      const int a0 = 2 * pr * bsz, b0 = a0 + bsz;
      T acc = T(0);
      if (LOWERM) {
        for (int k = c; k < bsz; ++k) acc = acc + sD[b0 + r][a0 + k] * sD[a0 + k][a0 + c];
      } else {
        for (int k = r; k < bsz; ++k) acc = acc + sD[a0 + r][a0 + k] * sD[a0 + k][b0 + c];
      }
      tmp[e] = acc;
    }
    __syncthreads();
    // phase 2: off-diagonal block = -inv(B) tmp (lower) or -tmp inv(B) (upper)
    for (int e = tid; e < cnt; e += NTH) {
      const int pr = e / (bsz * bsz), r = (e / bsz) % bsz, c = e % bsz;
      const int a0 = 2 * pr * bsz, b0 = a0 + bsz;
      const T* tp = tmp + (size_t)pr * bsz * bsz;
      T acc = T(0);
      if (LOWERM) {
        for (int k = 0; k <= r; ++k) acc = acc + sD[b0 + r][b0 + k] * tp[k * bsz + c];
        sD[b0 + r][a0 + c] = -acc;
      } else {
        for (int k = 0; k <= c; ++k) acc = acc + tp[r * bsz + k] * sD[b0 + k][b0 + c];
        sD[a0 + r][b0 + c] = -acc;
      }
    }
    __syncthreads();
  }
}

// QN: right-hand sides handled per launch (1, 4 or 16); QW of them per wave.
template <class T, int QN, bool LOWERM, bool UNIT, bool TRANS>
__global__ __launch_bounds__(256) void tri_chain_kernel(int n, int nrhs, const T* __restrict__ A, long lda,
                                                        T* __restrict__ X, long ldx, unsigned* ticket,
                                                        unsigned* flags, int* err) {
  constexpr int QW = QN < 4 ? QN : 4;
  constexpr int NWQ = QN / QW;  // waves taking part in the diagonal solve
  __shared__ T sD[TB][TB + 1];             // diagonal block of M (rows past n: identity)
  constexpr int QS = QN < 4 ? 4 : QN;      // >= 1024 values: the inversion's temporary
  __shared__ T sx[4][TB][QS];              // x blocks of a round; reused for the partial sums
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
  // this block's right-hand sides (rows only this workgroup writes), read
  // now so that the chain step after the last dependency does not wait for them
  T yb[QW];
#pragma unroll
  for (int u = 0; u < QW; ++u) {
    const int q = QW * (tid >> 6) + u;
    yb[u] = ((tid >> 6) < NWQ && (tid & (TB - 1)) < bs && q < nrhs) ? X[(long)(r0 + (tid & (TB - 1))) + (long)q * ldx]
                                                                 : T(0);
  }

  {
    T dv[TB * TB / 256];
#pragma unroll
    for (int it = 0; it < TB * TB / 256; ++it) {
      const int e = tid + it * 256;
      int i, c;
      if (TRANS) { c = e % TB; i = e / TB; } else { i = e % TB; c = e / TB; }
      // M's own triangle only (the other holds the other factor); rows and
      // columns past the matrix: identity
      const bool tri = LOWERM ? c <= i : c >= i;
      dv[it] = (i < bs && c < bs) ? (tri ? A[m_index<TRANS>(r0 + i, r0 + c, lda)] : T(0)) : (i == c ? T(1) : T(0));
    }
#pragma unroll
    for (int it = 0; it < TB * TB / 256; ++it) {
      const int e = tid + it * 256;
      int i, c;
      if (TRANS) { c = e % TB; i = e / TB; } else { i = e % TB; c = e / TB; }
      sD[i][c] = dv[it];
    }
  }
  __syncthreads();
  // the diagonal block's inverse, while the dependencies are outstanding
  invert_diag<T, LOWERM, UNIT>(sD, &sx[0][0][0]);

  const int i = tid & (TB - 1);
  const int g = tid >> 6;
  T acc[QN];
#pragma unroll
  for (int q = 0; q < QN; ++q) acc[q] = T(0);
  const int ndep = LOWERM ? bi : nblk - 1 - bi;
  bool failed = false;
  for (int d0 = 0; d0 < ndep;) {
    // rounds of up to DEPB over all but the last dependency, then the last alone
    const int nd = d0 < ndep - 1 ? min(DEPB, ndep - 1 - d0) : 1;
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
        ++spins;
        // a wait that gave up anywhere ends every other wait too
        if ((spins & 255u) == 1u && ld_dev(err) != 0) { failed = true; break; }
        if (spins > (1u << 22)) { st_dev(err, 1); failed = true; break; }
      }
    }
    __syncthreads();
    {
      // all of the round's x loads in flight at once, then into LDS
      constexpr int PER = (DEPB * TB * QN + 255) / 256;
      T xs[PER];
#pragma unroll
      for (int it = 0; it < PER; ++it) {
        const int e = tid + it * 256;
        const int c = e % TB, q = (e / TB) % QN, u = e / (TB * QN);
        const int j = LOWERM ? d0 + u : nblk - 1 - (d0 + u);
        xs[it] = (u < nd && q < nrhs && j * TB + c < n) ? ld_dev(&X[(long)(j * TB + c) + (long)q * ldx]) : T(0);
      }
#pragma unroll
      for (int it = 0; it < PER; ++it) {
        const int e = tid + it * 256;
        if (e < DEPB * TB * QN) sx[e / (TB * QN)][e % TB][(e / TB) % QN] = xs[it];
      }
    }
    __syncthreads();
    T p[QN];
#pragma unroll
    for (int q = 0; q < QN; ++q) p[q] = T(0);
#pragma unroll
    for (int u = 0; u < DEPB; ++u) {
      if (u < nd) {  // uniform: the last dependency's round holds one tile
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          const T m = mv[u][c];
#pragma unroll
          for (int q = 0; q < QN; ++q) p[q] = p[q] - m * sx[u][16 * g + c][q];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < QN; ++q) acc[q] = acc[q] + p[q];
    __syncthreads();
    d0 += nd;
  }

  // partial sums of the four column groups, then y = b + acc
  T* red = &sx[0][0][0];  // [4][TB][QN]
#pragma unroll
  for (int q = 0; q < QN; ++q) red[(g * TB + i) * QN + q] = acc[q];
  __syncthreads();
  // wave g solves right-hand sides QW*g .. QW*g+QW-1 of the diagonal block
  const int lane = i;
  T x[QW];
#pragma unroll
  for (int u = 0; u < QW; ++u) {
    const int q = QW * g + u;
    T y = T(0);
    if (g < NWQ && lane < bs && q < nrhs) {
      y = yb[u];
      y = y + (((red[(0 * TB + lane) * QN + q] + red[(1 * TB + lane) * QN + q]) + red[(2 * TB + lane) * QN + q]) +
               red[(3 * TB + lane) * QN + q]);
    }
    x[u] = y;
  }
  __syncthreads();  // every partial sum has been read before red is reused
  if (g < NWQ) {
    // x = inv(M_dd) y: the wave's y columns through LDS, then row `lane` of
    // the inverse times them
    T* sy = red + (size_t)g * TB * QN;
#pragma unroll
    for (int u = 0; u < QW; ++u) sy[lane * QW + u] = x[u];
    __builtin_amdgcn_wave_barrier();
    T xs2[QW];
#pragma unroll
    for (int u = 0; u < QW; ++u) xs2[u] = T(0);
#pragma unroll 8
    for (int k = 0; k < TB; ++k) {
      const T zik = sD[lane][k];
#pragma unroll
      for (int u = 0; u < QW; ++u) xs2[u] = xs2[u] + zik * sy[k * QW + u];
    }
#pragma unroll
    for (int u = 0; u < QW; ++u) x[u] = xs2[u];
#pragma unroll
    for (int u = 0; u < QW; ++u) {
      const int q = QW * g + u;
      if (lane < bs && q < nrhs) st_dev(&X[(long)(r0 + lane) + (long)q * ldx], x[u]);
    }
  }
  wait_vm();
  __syncthreads();
  if (tid == 0) st_dev(&flags[bi], 1u);
}

template <class T, int QN, bool LOWERM, bool UNIT, bool TRANS>
void launch_chain_q(int nblk, int n, int nq, const T* A, long lda, T* X, long ldx, unsigned* ctl, int* err,
                    hipStream_t s) {
  tri_chain_kernel<T, QN, LOWERM, UNIT, TRANS><<<nblk, 256, 0, s>>>(n, nq, A, lda, X, ldx, ctl, ctl + 1, err);
}

template <class T, bool LOWERM, bool UNIT, bool TRANS>
void launch_chain(int n, int nrhs, const T* A, long lda, T* X, long ldx, unsigned* ctl, int* err, hipStream_t s) {
  const int nblk = (n + TB - 1) / TB;
  for (int q0 = 0; q0 < nrhs; q0 += TQ) {
    const int nq = min(TQ, nrhs - q0);
    launch_fill_u32(ctl, (size_t)nblk + 1, 0u, s);  // ticket + flags (a kernel: graph-safe)
    T* Xq = X + (long)q0 * ldx;
    if (nq == 1) launch_chain_q<T, 1, LOWERM, UNIT, TRANS>(nblk, n, nq, A, lda, Xq, ldx, ctl, err, s);
    else if (nq <= 4) launch_chain_q<T, 4, LOWERM, UNIT, TRANS>(nblk, n, nq, A, lda, Xq, ldx, ctl, err, s);
    else launch_chain_q<T, 16, LOWERM, UNIT, TRANS>(nblk, n, nq, A, lda, Xq, ldx, ctl, err, s);
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
