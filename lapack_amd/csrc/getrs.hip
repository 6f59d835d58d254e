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

// One triangular sweep x = op(M)^-1 b (16 right-hand sides at most) as a
// chain workgroup plus one helper workgroup per 64-row block, so that the
// block-to-block chain runs inside one workgroup's registers and LDS instead
// of through memory.  Step u of the chain is block bi(u) (u-th from the top
// for a lower M, from the bottom for an upper one):
//     x_u = D_u^-1 (b_u - sum_{v<u} M_uv x_v)
//         = z_u - sum_{d=1..S} G_ud x_{u-d},   G_ud = D_u^-1 M_u,u-d,
//     z_u = D_u^-1 (b_u - sum_{v<u-S} M_uv x_v).
// Helper u (ticket u+1) inverts D_u, forms G_u1..G_uS into scratch while the
// chain is still far away, accumulates the far terms as the chain publishes
// x_v (v < u - S: flags xflag[v]), then writes z_u over b_u and raises
// ready[u].  The chain (ticket 0) keeps x_{u-S}..x_{u-1} in LDS: each step
// is one S-term product per row from registers, with the next step's z and
// G already loaded (fetched two steps ahead, once ready[] says so); it
// writes x_u over z_u and raises xflag[u] during the next step, once its
// stores have completed.  A workgroup only waits for tickets below its
// own, so every wait ends; every wait is bounded and polls *err, as in
// tri_chain_kernel.
constexpr int SWQ = 16;  // right-hand sides per sweep launch

template <bool TRANS>
__device__ __forceinline__ long sw_index(int i, int c, long lda) {
  return TRANS ? (long)c + (long)i * lda : (long)i + (long)c * lda;
}

template <class T, int S, bool LOWERM, bool UNIT, bool TRANS>
__global__ __launch_bounds__(512) void tri_sweep_kernel(int n, int nrhs, const T* __restrict__ A, long lda,
                                                        T* __restrict__ X, long ldx, unsigned* ctl,
                                                        T* __restrict__ Gbuf, int* err) {
  constexpr int QN = SWQ;
  constexpr int NG = 8;          // column groups: thread (i, g) owns columns 8g .. 8g+7 of a tile
  constexpr int CW = TB / NG;    // 8
  constexpr int QW = QN / NG;    // right-hand sides per thread in the final step: 2
  constexpr int LDD = TB + 1;
  // LDS, by role.  Helper: D^-1 [TB][LDD], then either a tile of M [TB][LDD]
  // and two x blocks [2][TB][QN], or the partial sums [NG][TB][QN].
  // Chain: the x ring [S][TB][QN] and the partial sums [NG][TB][QN].
  constexpr int HELP = TB * LDD + (TB * LDD + 2 * TB * QN > NG * TB * QN ? TB * LDD + 2 * TB * QN : NG * TB * QN);
  constexpr int CHAIN = S * TB * QN + NG * TB * QN;
  __shared__ T smem[HELP > CHAIN ? HELP : CHAIN];
  __shared__ int s_t, s_fail;
  const int tid = threadIdx.x;
  const int i = tid & (TB - 1);
  const int g = tid >> 6;
  const int nblk = (n + TB - 1) / TB;
  unsigned* ready = ctl + 1;
  unsigned* xflag = ctl + 1 + nblk;
  if (tid == 0) {
    s_t = (int)__hip_atomic_fetch_add(ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_fail = 0;
  }
  __syncthreads();
  const int t = s_t;
  if (t > nblk) return;
  auto blk = [&](int u) { return LOWERM ? u : nblk - 1 - u; };
  // bounded wait for a flag (one thread); a wait that gives up sets *err,
  // and every wait polls *err
  auto wait_flag = [&](const unsigned* f) -> bool {
    unsigned spins = 0;
    while (ld_dev(f) == 0u) {
      __builtin_amdgcn_s_sleep(1);
      ++spins;
      if ((spins & 255u) == 1u && ld_dev(err) != 0) return false;
      if (spins > (1u << 22)) {
        st_dev(err, 1);
        return false;
      }
    }
    return true;
  };

  if (t >= 1) {
    // ------------------------------------------------------------- helper
    T (*sD)[LDD] = reinterpret_cast<T (*)[LDD]>(smem);
    T (*sM)[LDD] = reinterpret_cast<T (*)[LDD]>(smem + TB * LDD);
    T* sx = smem + 2 * TB * LDD;   // [2][TB][QN]
    T* red = smem + TB * LDD;      // [NG][TB][QN], after the tiles are done with
    const int u = t - 1;
    const int bi = blk(u);
    const int r0 = bi * TB;
    const int bs = min(TB, n - r0);
    // own right-hand sides: thread (i, g) holds b[i][QW g .. QW g + QW-1]
    T yb[QW];
#pragma unroll
    for (int q = 0; q < QW; ++q) {
      const int qq = QW * g + q;
      yb[q] = (i < bs && qq < nrhs) ? X[(long)(r0 + i) + (long)qq * ldx] : T(0);
    }
    // D_u (rows and columns past n: identity), inverted in sD
    for (int e = tid; e < TB * TB; e += 512) {
      int r, c;
      if (TRANS) { c = e % TB; r = e / TB; } else { r = e % TB; c = e / TB; }
      const bool tri = LOWERM ? c <= r : c >= r;
      sD[r][c] = (r < bs && c < bs) ? (tri ? A[sw_index<TRANS>(r0 + r, r0 + c, lda)] : T(0)) : (r == c ? T(1) : T(0));
    }
    __syncthreads();
    invert_diag<T, LOWERM, UNIT, 512>(sD, sx);
    // G_ud = D_u^-1 M_u,u-d for d = 1 .. min(S, u), row-major TB x TB in
    // Gbuf[(u*S + d-1)*TB*TB]: thread (i, g) forms row i, columns 8g .. 8g+7
    for (int d = 1; d <= min(S, u); ++d) {
      const int c0 = blk(u - d) * TB;
      const int cs = min(TB, n - c0);
      for (int e = tid; e < TB * TB; e += 512) {
        int r, c;
        if (TRANS) { c = e % TB; r = e / TB; } else { r = e % TB; c = e / TB; }
        sM[r][c] = (r < bs && c < cs) ? A[sw_index<TRANS>(r0 + r, c0 + c, lda)] : T(0);
      }
      __syncthreads();
      T gacc[CW];
#pragma unroll
      for (int c = 0; c < CW; ++c) gacc[c] = T(0);
#pragma unroll 4
      for (int k = 0; k < TB; ++k) {
        const T dik = sD[i][k];
#pragma unroll
        for (int c = 0; c < CW; ++c) gacc[c] = gacc[c] + dik * sM[k][CW * g + c];
      }
      T* gp = Gbuf + ((size_t)u * S + (d - 1)) * TB * TB + (size_t)i * TB + CW * g;
#pragma unroll
      for (int c = 0; c < CW; ++c) gp[c] = gacc[c];
      __syncthreads();  // sM is reloaded next
    }
    // far terms v < u - S, two dependency blocks per round, their M values
    // issued before the wait
    T acc[QN];
#pragma unroll
    for (int q = 0; q < QN; ++q) acc[q] = T(0);
    const int nfar = u - S;
    bool failed = false;
    for (int v0 = 0; v0 < nfar; v0 += 2) {
      const int nd = min(2, nfar - v0);
      T mv[2][CW];
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        const int c0 = blk(v0 + w) * TB + CW * g;
#pragma unroll
        for (int c = 0; c < CW; ++c) {
          T val = T(0);
          if (w < nd && i < bs && c0 + c < n) val = A[sw_index<TRANS>(r0 + i, c0 + c, lda)];
          mv[w][c] = val;
        }
      }
      if (tid < nd && !failed) {
        if (!wait_flag(&xflag[v0 + tid])) s_fail = 1;
      }
      __syncthreads();
      failed = failed || s_fail != 0;
      {
        constexpr int PER = (2 * TB * QN) / 512;
        T xs[PER];
#pragma unroll
        for (int it = 0; it < PER; ++it) {
          const int e = tid + it * 512;
          const int c = e % TB, q = (e / TB) % QN, w = e / (TB * QN);
          const int row = blk(v0 + w) * TB + c;
          xs[it] = (w < nd && q < nrhs && row < n) ? ld_dev(&X[(long)row + (long)q * ldx]) : T(0);
        }
#pragma unroll
        for (int it = 0; it < PER; ++it) {
          const int e = tid + it * 512;
          sx[((e / (TB * QN)) * TB + e % TB) * QN + (e / TB) % QN] = xs[it];
        }
      }
      __syncthreads();
      T p[QN];
#pragma unroll
      for (int q = 0; q < QN; ++q) p[q] = T(0);
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        if (w < nd) {
#pragma unroll
          for (int c = 0; c < CW; ++c) {
            const T m = mv[w][c];
            const T* xr = sx + (w * TB + CW * g + c) * QN;
#pragma unroll
            for (int q = 0; q < QN; ++q) p[q] = p[q] - m * xr[q];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < QN; ++q) acc[q] = acc[q] + p[q];
      __syncthreads();
    }
    // y = b + the column groups' partial sums (in group order), z = D^-1 y
    // (thread (i, g): z[i][QW g ..]), written over b
#pragma unroll
    for (int q = 0; q < QN; ++q) red[(g * TB + i) * QN + q] = acc[q];
    __syncthreads();
    T y[QW];
#pragma unroll
    for (int q = 0; q < QW; ++q) {
      const int qq = QW * g + q;
      T sum = red[(0 * TB + i) * QN + qq];
#pragma unroll
      for (int gg = 1; gg < NG; ++gg) sum = sum + red[(gg * TB + i) * QN + qq];
      y[q] = yb[q] + sum;
    }
    __syncthreads();
    T* sy = red;  // [TB][QN]
#pragma unroll
    for (int q = 0; q < QW; ++q) sy[i * QN + QW * g + q] = y[q];
    __syncthreads();
    T z[QW];
#pragma unroll
    for (int q = 0; q < QW; ++q) z[q] = T(0);
#pragma unroll 8
    for (int k = 0; k < TB; ++k) {
      const T dik = sD[i][k];
#pragma unroll
      for (int q = 0; q < QW; ++q) z[q] = z[q] + dik * sy[k * QN + QW * g + q];
    }
    if (!failed) {
#pragma unroll
      for (int q = 0; q < QW; ++q) {
        const int qq = QW * g + q;
        if (i < bs && qq < nrhs) st_dev(&X[(long)(r0 + i) + (long)qq * ldx], z[q]);
      }
    }
    wait_vm();  // z and G stored
    __syncthreads();
    if (tid == 0) st_dev(&ready[u], failed ? 2u : 1u);
    return;
  }

  // --------------------------------------------------------------- chain
  T* ring = smem;                // [S][TB][QN]: x_{u-S} .. x_{u-1}, slot u % S
  T* red = smem + S * TB * QN;   // [NG][TB][QN]
  T gv[2][S][CW];  // G_u,d[i][8g .. 8g+7], two steps in flight
  T zv[2][QW];     // z_u[i][QW g ..]
  bool failed = false;
  // step u's z and G into register set h (after ready[u])
  auto fetch = [&](int u, int h) {
    const int r0 = blk(u) * TB;
    const int bs = min(TB, n - r0);
#pragma unroll
    for (int q = 0; q < QW; ++q) {
      const int qq = QW * g + q;
      zv[h][q] = (i < bs && qq < nrhs) ? ld_dev(&X[(long)(r0 + i) + (long)qq * ldx]) : T(0);
    }
#pragma unroll
    for (int d = 1; d <= S; ++d) {
      const T* gp = Gbuf + ((size_t)u * S + (d - 1)) * TB * TB + (size_t)i * TB + CW * g;
#pragma unroll
      for (int c = 0; c < CW; ++c) gv[h][d - 1][c] = (d <= u) ? ld_dev(&gp[c]) : T(0);
    }
  };
  // one thread waits for ready[u] (u < nblk); the result through LDS
  auto await_ready = [&](int u) {
    if (tid == 0 && !failed && u < nblk) {
      if (!wait_flag(&ready[u])) s_fail = 1;
      else if (ld_dev(&ready[u]) != 1u) s_fail = 1;  // the helper's own wait gave up
    }
    __syncthreads();
    failed = failed || s_fail != 0;
  };
  for (int e = tid; e < S * TB * QN; e += 512) ring[e] = T(0);
  await_ready(0);
  if (!failed) fetch(0, 0);
  await_ready(1);
  if (!failed && nblk > 1) fetch(1, 1);
  auto step = [&](int u, int h) {
    const int r0 = blk(u) * TB;
    const int bs = min(TB, n - r0);
    // p = sum_d G_u,d x_{u-d} over this thread's CW columns
    T p[QN];
#pragma unroll
    for (int q = 0; q < QN; ++q) p[q] = T(0);
#pragma unroll
    for (int d = 1; d <= S; ++d) {
      const T* xs = ring + (size_t)(((u - d) % S + S) % S) * TB * QN;
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        const T gval = gv[h][d - 1][c];
        const T* xr = xs + (CW * g + c) * QN;
#pragma unroll
        for (int q = 0; q < QN; ++q) p[q] = p[q] + gval * xr[q];
      }
    }
#pragma unroll
    for (int q = 0; q < QN; ++q) red[(g * TB + i) * QN + q] = p[q];
    __syncthreads();
    // the previous step's x stores have completed (every wave waited for
    // younger loads since): publish it
    if (tid == 0 && u > 0 && !failed) st_dev(&xflag[u - 1], 1u);
    T x[QW];
#pragma unroll
    for (int q = 0; q < QW; ++q) {
      const int qq = QW * g + q;
      T sum = red[(0 * TB + i) * QN + qq];
#pragma unroll
      for (int gg = 1; gg < NG; ++gg) sum = sum + red[(gg * TB + i) * QN + qq];
      x[q] = zv[h][q] - sum;
    }
    // x_u over slot u % S (x_{u-S}, read above before the barrier)
    T* xo = ring + (size_t)(u % S) * TB * QN;
#pragma unroll
    for (int q = 0; q < QW; ++q) xo[i * QN + QW * g + q] = (i < bs) ? x[q] : T(0);
    if (!failed) {
#pragma unroll
      for (int q = 0; q < QW; ++q) {
        const int qq = QW * g + q;
        if (i < bs && qq < nrhs) st_dev(&X[(long)(r0 + i) + (long)qq * ldx], x[q]);
      }
    }
    __syncthreads();  // ring and red settled
    // step u + 2's z and G into this set
    await_ready(u + 2);
    if (!failed && u + 2 < nblk) fetch(u + 2, h);
  };
#pragma nounroll
  for (int u = 0; u < nblk; u += 2) {
    step(u, 0);
    if (u + 1 < nblk) step(u + 1, 1);
  }
  wait_vm();
  __syncthreads();
  if (tid == 0 && !failed) st_dev(&xflag[nblk - 1], 1u);
  if (tid == 0 && failed) st_dev(err, 1);
}

template <class T, int S, bool LOWERM, bool UNIT, bool TRANS>
void launch_sweep_t(int n, int nrhs, const T* A, long lda, T* X, long ldx, unsigned* ctl, T* G, int* err,
                    hipStream_t s) {
  const int nblk = (n + TB - 1) / TB;
  for (int q0 = 0; q0 < nrhs; q0 += SWQ) {
    const int nq = min(SWQ, nrhs - q0);
    launch_fill_u32(ctl, (size_t)2 * nblk + 1, 0u, s);  // ticket, ready, xflag (a kernel: graph-safe)
    tri_sweep_kernel<T, S, LOWERM, UNIT, TRANS><<<nblk + 1, 512, 0, s>>>(n, nq, A, lda, X + (long)q0 * ldx, ldx, ctl,
                                                                         G, err);
  }
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

constexpr int SWEEP_S = 3;  // near blocks the sweep's chain workgroup takes itself
size_t tri_sweep_ctl_words(int n) { return (size_t)2 * ((n + TB - 1) / TB) + 1; }
size_t tri_sweep_g_elems(int n) { return (size_t)((n + TB - 1) / TB) * SWEEP_S * TB * TB; }

// which: 0 = L (unit lower), 1 = U (upper), 2 = U^T, 3 = L^T (as launch_tri_chain).
template <class T>
void launch_tri_sweep(int which, int n, int nrhs, const T* A, long lda, T* X, long ldx, unsigned* ctl, T* G,
                      int* err, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  switch (which) {
    case 0: launch_sweep_t<T, SWEEP_S, true, true, false>(n, nrhs, A, lda, X, ldx, ctl, G, err, s); break;
    case 1: launch_sweep_t<T, SWEEP_S, false, false, false>(n, nrhs, A, lda, X, ldx, ctl, G, err, s); break;
    case 2: launch_sweep_t<T, SWEEP_S, true, false, true>(n, nrhs, A, lda, X, ldx, ctl, G, err, s); break;
    case 3: launch_sweep_t<T, SWEEP_S, false, true, true>(n, nrhs, A, lda, X, ldx, ctl, G, err, s); break;
  }
}
template void launch_tri_sweep<double>(int, int, int, const double*, long, double*, long, unsigned*, double*, int*,
                                       hipStream_t);
template void launch_tri_sweep<float>(int, int, int, const float*, long, float*, long, unsigned*, float*, int*,
                                      hipStream_t);

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
