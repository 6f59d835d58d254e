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

#include <algorithm>

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

// The sentinel that marks a value of x not yet written: a quiet NaN with a
// fixed payload.  A computed value with exactly these bits (only a NaN
// carrying this payload from the input) is stored as the NaN `alt` instead.
template <class T> struct Sent;
template <> struct Sent<double> {
  using U = unsigned long long;
  static constexpr U bits = 0x7FFC00005E471E57ull, alt = 0x7FFC00005E471E56ull;
};
template <> struct Sent<float> {
  using U = unsigned;
  static constexpr U bits = 0x7FC5E471u, alt = 0x7FC5E470u;
};
template <class T>
__device__ __forceinline__ bool is_sent(T v) {
  return __builtin_bit_cast(typename Sent<T>::U, v) == Sent<T>::bits;
}
template <class T>
__device__ __forceinline__ T canon(T v) {
  return is_sent(v) ? __builtin_bit_cast(T, Sent<T>::alt) : v;
}

constexpr unsigned SPIN_LIMIT = 1u << 22;

// FP64 / FP32 16x16x4 matrix-core product on gfx950: lane l holds A(l%16,
// l/16) and B(l/16, l%16) of a k-step; accumulator register r of lane l
// holds D(row(l/16, r), l%16).
template <class T> struct TriMfma;
template <> struct TriMfma<double> {
  typedef la_d4 acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct TriMfma<float> {
  typedef la_f4 acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// QN: right-hand sides handled per launch (1, 4 or 16).  Reads the
// right-hand sides from Xin, writes x to X (filled with the sentinel
// beforehand; the two may not overlap).
template <class T, int QN, bool LOWERM, bool UNIT, bool TRANS>
__global__ __launch_bounds__(256) void tri_chain_kernel(int n, int nrhs, const T* __restrict__ A, long lda,
                                                        const T* __restrict__ Xin, long ldxi, T* __restrict__ X,
                                                        long ldx, unsigned* ticket, int* err) {
  // MF: 16 right-hand sides, products on the matrix cores, wave g owning
  // rows 16g .. 16g+15 of the block; otherwise products on the vector units
  constexpr bool MF = QN == 16;
  typedef TriMfma<T> MM;
  typedef typename MM::acc_t acc_t;
  constexpr int XS = QN + 1;  // row stride of the [row][rhs] stagings
  constexpr int NBUF_MF = DEPB * TB * XS;
  constexpr int NBUF_V = (DEPB * TB * QN > 4 * TB * QN ? DEPB * TB * QN : 4 * TB * QN);
  constexpr int NB0 = MF ? NBUF_MF : NBUF_V;
  constexpr int NBUF = NB0 > TB * TB / 4 ? NB0 : TB * TB / 4;
  constexpr int WR = MF ? 1 : TB;  // sW rows: the vector form keeps M_last in LDS
  __shared__ T sD[TB][TB + 1];     // D, then inv(D) transposed: sD[k][i] = inv(D)(i, k)
  __shared__ T sW[WR][TB + 1];     // vector form: M_last transposed, sW[k][i] = M(i, k)
  __shared__ T sbuf[NBUF];         // inversion temporary, a round's x blocks, the partial sums
  __shared__ T sy[TB][XS];         // y, then y - M_last x_last
  __shared__ T sxl[TB][XS];        // x of the last dependency
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
  const int i = tid & (TB - 1);
  const int g = tid >> 6;
  const int lane = i;
  const int li = lane & 15, lk = lane >> 4;
  const int ndep = LOWERM ? bi : nblk - 1 - bi;
  const int jl = LOWERM ? bi - 1 : bi + 1;  // the last dependency (finished most recently)
  auto dep = [&](int d) { return LOWERM ? d : nblk - 1 - d; };

  // this block's right-hand sides, in the layout the far sum ends in
  constexpr int NYB = MF ? 4 : 1;
  T yb[NYB];
  if (MF) {
#pragma unroll
    for (int r = 0; r < NYB; ++r) {
      const int row = 16 * g + MM::row(lk, r);
      yb[r] = (row < bs && li < nrhs) ? Xin[(long)(r0 + row) + (long)li * ldxi] : T(0);
    }
  } else {
    yb[0] = (g < nrhs && i < bs) ? Xin[(long)(r0 + i) + (long)g * ldxi] : T(0);
  }
  // M_last, the last dependency's tile: matrix-core A operands in registers
  // (MF) or transposed into sW
  T mL[MF ? 16 : 1];
  if (ndep > 0) {
    const int c0 = jl * TB;
    if (MF) {
#pragma unroll
      for (int s4 = 0; s4 < 16; ++s4) {
        const int r = 16 * g + li, c = 4 * s4 + lk;
        mL[s4] = (r < bs && c0 + c < n) ? A[m_index<TRANS>(r0 + r, c0 + c, lda)] : T(0);
      }
    } else {
#pragma unroll 4
      for (int it = 0; it < TB * TB / 256; ++it) {
        const int e = tid + it * 256;
        int r, c;
        if (TRANS) { c = e % TB; r = e / TB; } else { r = e % TB; c = e / TB; }
        sW[c][r] = (r < bs && c0 + c < n) ? A[m_index<TRANS>(r0 + r, c0 + c, lda)] : T(0);
      }
    }
  }
  {
    T dv[TB * TB / 256];
#pragma unroll
    for (int it = 0; it < TB * TB / 256; ++it) {
      const int e = tid + it * 256;
      int r, c;
      if (TRANS) { c = e % TB; r = e / TB; } else { r = e % TB; c = e / TB; }
      // M's own triangle only (the other holds the other factor); rows and
      // columns past the matrix: identity
      const bool tri = LOWERM ? c <= r : c >= r;
      dv[it] = (r < bs && c < bs) ? (tri ? A[m_index<TRANS>(r0 + r, r0 + c, lda)] : T(0)) : (r == c ? T(1) : T(0));
    }
#pragma unroll
    for (int it = 0; it < TB * TB / 256; ++it) {
      const int e = tid + it * 256;
      int r, c;
      if (TRANS) { c = e % TB; r = e / TB; } else { r = e % TB; c = e / TB; }
      sD[r][c] = dv[it];
    }
  }
  __syncthreads();
  invert_diag<T, LOWERM, UNIT>(sD, sbuf);  // ends with a barrier
  {
    // transpose in place, so that lane i reads inv(D)(i, k) at sD[k][i]
    T a[TB * TB / 256], b[TB * TB / 256];
#pragma unroll
    for (int it = 0; it < TB * TB / 256; ++it) {
      const int e = tid + it * 256;
      const int r = e / TB, c = e % TB;
      a[it] = sD[r][c];
      b[it] = sD[c][r];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < TB * TB / 256; ++it) {
      const int e = tid + it * 256;
      const int r = e / TB, c = e % TB;
      if (r < c) { sD[r][c] = b[it]; sD[c][r] = a[it]; }
    }
  }

  // far dependencies (all but the last), in the order they finish, DEPB per
  // round; their M values are read one round ahead, so only the x blocks are
  // waited for
  const int nfar = ndep > 0 ? ndep - 1 : 0;
  bool failed = false;
  acc_t facc = acc_t{};       // MF: the wave's 16 x 16 block of sum M x
  T vacc[MF ? 1 : QN];        // vector form: row i, columns 16g .. 16g+15 of each tile
#pragma unroll
  for (int q = 0; q < (MF ? 1 : QN); ++q) vacc[q] = T(0);
  // M values of a round: MF, lane (li, lk) of wave g holds M(16g+li, 4s+lk)
  // for the 16 k-steps s; vector form, thread (i, g) holds M(i, 16g+c)
  auto load_m = [&](T(&m)[DEPB][16], int d0) {
#pragma unroll
    for (int u = 0; u < DEPB; ++u) {
      const int cd = dep(d0 + u) * TB;
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const int r = MF ? 16 * g + li : i;
        const int cc = MF ? 4 * c + lk : 16 * g + c;
        T v = T(0);
        if (d0 + u < nfar && r < bs && cd + cc < n) v = A[m_index<TRANS>(r0 + r, cd + cc, lda)];
        m[u][c] = v;
      }
    }
  };
  auto round = [&](const T(&m)[DEPB][16], int d0) {
    const int nd = min(DEPB, nfar - d0);
    constexpr int PER = (DEPB * TB * QN + 255) / 256;
    T xs[PER];
    unsigned spins = 0;
    // first one word per dependency, so that waiting blocks re-read a few
    // words, not whole x blocks
    {
      int bad = 0;
      if (tid < nd) {
        const T* w = &X[(long)(dep(d0 + tid) * TB)];
        while (is_sent(ld_dev(w))) {
          __builtin_amdgcn_s_sleep(1);
          ++spins;
          if ((spins & 255u) == 0u && ld_dev(err) != 0) { bad = 1; break; }
          if (spins > SPIN_LIMIT) { st_dev(err, 1); bad = 1; break; }
        }
      }
      if (__syncthreads_or(bad)) { failed = true; return; }
      spins = 0;
    }
    for (;;) {
      int miss = 0;
#pragma unroll
      for (int it = 0; it < PER; ++it) {
        const int e = tid + it * 256;
        const int c = e % TB, q = (e / TB) % QN, u = e / (TB * QN);
        const int row = dep(d0 + u) * TB + c;
        const bool valid = u < nd && q < nrhs && row < n;
        xs[it] = valid ? ld_dev(&X[(long)row + (long)q * ldx]) : T(0);
        miss |= (valid && is_sent(xs[it])) ? 1 : 0;
      }
      if (!__syncthreads_or(miss)) break;
      __builtin_amdgcn_s_sleep(1);
      ++spins;
      // a wait that gave up anywhere ends every other wait too
      if ((spins & 255u) == 0u && __syncthreads_or(tid == 0 && ld_dev(err) != 0)) { failed = true; break; }
      if (spins > SPIN_LIMIT) {
        if (tid == 0) st_dev(err, 1);
        failed = true;
        break;
      }
    }
    if (failed) return;
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int e = tid + it * 256;
      const int c = e % TB, q = (e / TB) % QN, u = e / (TB * QN);
      if (e < DEPB * TB * QN) {
        if (MF) sbuf[(u * TB + c) * XS + q] = xs[it];  // [u][k][q]
        else sbuf[e] = xs[it];                          // [u][q][c]
      }
    }
    __syncthreads();
    if (MF) {
      // the round's products summed on their own, then added to the
      // running sum (no sum runs over more than one round's 128 terms or
      // one term per round)
      acc_t racc = acc_t{};
#pragma unroll
      for (int u = 0; u < DEPB; ++u) {
        if (u < nd) {
#pragma unroll
          for (int s4 = 0; s4 < 16; ++s4) racc = MM::op(m[u][s4], sbuf[(u * TB + 4 * s4 + lk) * XS + li], racc);
        }
      }
      facc = facc + racc;
    } else {
      T p[MF ? 1 : QN];
#pragma unroll
      for (int q = 0; q < (MF ? 1 : QN); ++q) p[q] = T(0);
#pragma unroll
      for (int u = 0; u < DEPB; ++u) {
        if (u < nd) {
#pragma unroll
          for (int c = 0; c < 16; ++c) {
            const T mv = m[u][c];
#pragma unroll
            for (int q = 0; q < (MF ? 1 : QN); ++q) p[q] = p[q] + mv * sbuf[(u * QN + q) * TB + 16 * g + c];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < (MF ? 1 : QN); ++q) vacc[q] = vacc[q] + p[q];
    }
    __syncthreads();
  };
  {
    T mA[DEPB][16], mB[DEPB][16];
    load_m(mA, 0);
    for (int d0 = 0; d0 < nfar && !failed; d0 += 2 * DEPB) {
      load_m(mB, d0 + DEPB);
      round(mA, d0);
      if (d0 + DEPB < nfar && !failed) {
        load_m(mA, d0 + 2 * DEPB);
        round(mB, d0 + DEPB);
      }
    }
  }
  if (failed) return;  // uniform

  // y = b - (far sum) into sy
  if (MF) {
#pragma unroll
    for (int r = 0; r < 4; ++r) sy[16 * g + MM::row(lk, r)][li] = yb[r] - facc[r];
  } else {
#pragma unroll
    for (int q = 0; q < (MF ? 1 : QN); ++q) sbuf[(g * TB + i) * QN + q] = vacc[q];
    __syncthreads();
    if (g < QN)
      sy[i][g] = yb[0] - (((sbuf[(0 * TB + i) * QN + g] + sbuf[(1 * TB + i) * QN + g]) + sbuf[(2 * TB + i) * QN + g]) +
                          sbuf[(3 * TB + i) * QN + g]);
  }
  __syncthreads();

  if (ndep > 0) {
    // the chain step: wait for x(block jl), then y -= M_last x(jl)
    const int bsj = min(TB, n - jl * TB);
    constexpr int PL = (TB * QN + 255) / 256;
    T v[PL];
    unsigned spins = 0;
    for (;;) {
      int miss = 0;
#pragma unroll
      for (int it = 0; it < PL; ++it) {
        const int e = tid + it * 256;
        const int k = e % TB, q = e / TB;
        const bool valid = e < TB * QN && k < bsj && q < nrhs;
        v[it] = valid ? ld_dev(&X[(long)(jl * TB + k) + (long)q * ldx]) : T(0);
        miss |= (valid && is_sent(v[it])) ? 1 : 0;
      }
      if (!__syncthreads_or(miss)) break;
      __builtin_amdgcn_s_sleep(1);
      ++spins;
      if ((spins & 255u) == 0u && __syncthreads_or(tid == 0 && ld_dev(err) != 0)) return;
      if (spins > SPIN_LIMIT) {
        if (tid == 0) st_dev(err, 1);
        return;
      }
    }
#pragma unroll
    for (int it = 0; it < PL; ++it) {
      const int e = tid + it * 256;
      if (e < TB * QN) sxl[e % TB][e / TB] = v[it];
    }
    __syncthreads();
    if (MF) {
      acc_t a2 = acc_t{};
#pragma unroll
      for (int s4 = 0; s4 < 16; ++s4) a2 = MM::op(mL[s4], sxl[4 * s4 + lk][li], a2);
#pragma unroll
      for (int r = 0; r < 4; ++r) sy[16 * g + MM::row(lk, r)][li] -= a2[r];
    } else if (g < QN) {
      T p0 = T(0), p1 = T(0), p2 = T(0), p3 = T(0);
#pragma unroll 4
      for (int k = 0; k < TB; k += 4) {
        p0 = p0 + sW[k][i] * sxl[k][g];
        p1 = p1 + sW[k + 1][i] * sxl[k + 1][g];
        p2 = p2 + sW[k + 2][i] * sxl[k + 2][g];
        p3 = p3 + sW[k + 3][i] * sxl[k + 3][g];
      }
      sy[i][g] -= (p0 + p1) + (p2 + p3);
    }
    __syncthreads();
  }
  // x = inv(D) y
  if (MF) {
    acc_t a3 = acc_t{};
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4) a3 = MM::op(sD[4 * s4 + lk][16 * g + li], sy[4 * s4 + lk][li], a3);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * g + MM::row(lk, r);
      if (row < bs && li < nrhs) st_dev(&X[(long)(r0 + row) + (long)li * ldx], canon(a3[r]));
    }
  } else if (g < QN) {
    T p0 = T(0), p1 = T(0), p2 = T(0), p3 = T(0);
#pragma unroll 4
    for (int k = 0; k < TB; k += 4) {
      p0 = p0 + sD[k][i] * sy[k][g];
      p1 = p1 + sD[k + 1][i] * sy[k + 1][g];
      p2 = p2 + sD[k + 2][i] * sy[k + 2][g];
      p3 = p3 + sD[k + 3][i] * sy[k + 3][g];
    }
    if (i < bs && g < nrhs) st_dev(&X[(long)(r0 + i) + (long)g * ldx], canon((p0 + p1) + (p2 + p3)));
  }
}

// X(0:n, 0:nrhs) = the sentinel (a kernel: graph-safe)
template <class T>
__global__ void fill_sentinel_kernel(int n, int nrhs, T* X, long ldx) {
  const long total = (long)n * nrhs;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
    X[(e % n) + (e / n) * ldx] = __builtin_bit_cast(T, Sent<T>::bits);
}

template <class T, int QN, bool LOWERM, bool UNIT, bool TRANS>
void launch_chain_q(int nblk, int n, int nq, const T* A, long lda, const T* Xin, long ldxi, T* X, long ldx,
                    unsigned* ticket, int* err, hipStream_t s) {
  tri_chain_kernel<T, QN, LOWERM, UNIT, TRANS><<<nblk, 256, 0, s>>>(n, nq, A, lda, Xin, ldxi, X, ldx, ticket, err);
}

template <class T, bool LOWERM, bool UNIT, bool TRANS>
void launch_chain(int n, int nrhs, const T* A, long lda, const T* Xin, long ldxi, T* X, long ldx, unsigned* ctl,
                  int* err, hipStream_t s) {
  const int nblk = (n + TB - 1) / TB;
  const int nsl = (nrhs + TQ - 1) / TQ;
  launch_fill_u32(ctl, (size_t)nsl, 0u, s);  // one ticket per slice of right-hand sides
  const long total = (long)n * nrhs;
  const int fb = (int)std::min<long>((total + 255) / 256, 2048);
  fill_sentinel_kernel<T><<<fb, 256, 0, s>>>(n, nrhs, X, ldx);
  for (int sl = 0; sl < nsl; ++sl) {
    const int q0 = sl * TQ;
    const int nq = min(TQ, nrhs - q0);
    const T* Xi = Xin + (long)q0 * ldxi;
    T* Xq = X + (long)q0 * ldx;
    unsigned* tk = ctl + sl;
    if (nq == 1) launch_chain_q<T, 1, LOWERM, UNIT, TRANS>(nblk, n, nq, A, lda, Xi, ldxi, Xq, ldx, tk, err, s);
    else if (nq <= 4) launch_chain_q<T, 4, LOWERM, UNIT, TRANS>(nblk, n, nq, A, lda, Xi, ldxi, Xq, ldx, tk, err, s);
    else launch_chain_q<T, 16, LOWERM, UNIT, TRANS>(nblk, n, nq, A, lda, Xi, ldxi, Xq, ldx, tk, err, s);
  }
}

}  // namespace

size_t tri_chain_ctl_words(int nrhs) { return (size_t)std::max(1, (nrhs + TQ - 1) / TQ); }

// which: 0 = L (unit lower), 1 = U (upper), 2 = U^T, 3 = L^T.
template <class T>
void launch_tri_chain(int which, int n, int nrhs, const T* A, long lda, const T* Xin, long ldxi, T* X, long ldx,
                      unsigned* ctl, int* err, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return;
  switch (which) {
    case 0: launch_chain<T, true, true, false>(n, nrhs, A, lda, Xin, ldxi, X, ldx, ctl, err, s); break;
    case 1: launch_chain<T, false, false, false>(n, nrhs, A, lda, Xin, ldxi, X, ldx, ctl, err, s); break;
    case 2: launch_chain<T, true, false, true>(n, nrhs, A, lda, Xin, ldxi, X, ldx, ctl, err, s); break;
    case 3: launch_chain<T, false, true, true>(n, nrhs, A, lda, Xin, ldxi, X, ldx, ctl, err, s); break;
  }
}
template void launch_tri_chain<double>(int, int, int, const double*, long, const double*, long, double*, long,
                                       unsigned*, int*, hipStream_t);
template void launch_tri_chain<float>(int, int, int, const float*, long, const float*, long, float*, long,
                                      unsigned*, int*, hipStream_t);

}  // namespace la
