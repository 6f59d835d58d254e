// panel.hip -- leaf factorization of a tall panel (the unblocked core of
// DGETRF2, SRC/dgetrf2.f:179-214 applied column by column).
//
// The panel is A(j0:m-1, j0:j0+w-1) of the matrix being factored, held in the
// transposed workspace T = A^T, so matrix row r of the panel is T column r and
// its w values are contiguous.  Three forms, chosen by the panel height h:
//   * h <= 1024: one workgroup keeps the whole leaf in registers
//     (getf2_panel_kernel);
//   * 1024 < h <= 32768: G = h/512 workgroups keep it in registers and agree
//     on each column's pivot through a flag-in-data exchange in device memory
//     (getf2_panel_w8_kernel; needs its G workgroups resident together);
//   * taller panels, and every leaf of a factorization being redone after a
//     bounded wait gave up: one column per two launches over the leaf in HBM
//     (tall_*_kernel), no workgroup waits for another.
// Per column k every form does:
//   1. pivot search under the IDAMAX rule of BLAS/SRC/idamax.f:101-107 (first
//      index of max |x|, strict '>', a NaN at the head of the column wins, NaNs
//      elsewhere never do).  Each candidate gets a 64-bit key that orders
//      exactly like that rule: bits(|x|)+1 (IEEE bit patterns of non-negative
//      doubles sort like their values), row k's NaN -> bits(+inf)+1, other
//      NaNs -> 1, rows outside the active range -> 0; ties go to the lowest
//      row.  A wave takes the max key with DPP row shifts/broadcasts and the
//      lowest lane holding it with a ballot;
//   2. interchange (row k takes the pivot row, the pivot row takes the old row
//      k: SRC/dgetrf2.f:196-200), scaling by the reciprocal of the pivot, or
//      division when |pivot| < sfmin (SRC/dgetrf2.f:204-210), and the rank-1
//      update of the remaining columns (SRC/dgetrf2.f:245-246 one column at a
//      time), applied also for a zero pivot as the reference's DGEMM has no
//      zero test.  Every form does the same operations on each element, so
//      the factors do not depend on the form.
// IPIV receives absolute 1-based rows; INFO the first exact zero pivot
// (SRC/dgetrf2.f:192,212-214).  Every wait is bounded: a timeout sets *err,
// later waits give up at once, and the kernel always runs to its end.
#include "la_common.h"

#include <cstdlib>
#include <cstring>

namespace la {
namespace {

constexpr int GMAX = 64;            // workgroups of a register-resident leaf
constexpr int WMAX = 64;            // widest leaf

template <class T> struct Sfmin;
template <> struct Sfmin<double> { static constexpr double v = 0x1p-1022; };
template <> struct Sfmin<float> { static constexpr float v = 0x1p-126f; };

typedef unsigned long long u64;

// XCDs a rank-mode panel of G workgroups occupies (32 CUs each).
__host__ __device__ __forceinline__ unsigned panel_xcds(int G) { return (unsigned)((G + 31) / 32); }

// Pivot key: see the header comment.
template <class T>
__device__ __forceinline__ u64 pivot_key(T x, bool is_k) {
  const double a = fabs((double)x);
  if (a != a) return is_k ? 0x7FF0000000000001ull : 1ull;
  return (u64)__double_as_longlong(a) + 1ull;
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ u64 dpp_u64(u64 v) {
  const int lo = (int)(unsigned)(v & 0xffffffffull), hi = (int)(unsigned)(v >> 32);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xf, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xf, false);
  return ((u64)(unsigned)hi2 << 32) | (u64)(unsigned)lo2;
}
__device__ __forceinline__ u64 umax64(u64 a, u64 b) { return a > b ? a : b; }

// Max of a 64-bit key over the wave (uniform result): row_shr 1,2,4,8 leave
// each 16-lane row's max in its lane 15; row_bcast 15 and 31 carry it to lane 63.
__device__ __forceinline__ u64 wave_max_u64(u64 v) {
  v = umax64(v, dpp_u64<0x111, 0xf>(v));
  v = umax64(v, dpp_u64<0x112, 0xf>(v));
  v = umax64(v, dpp_u64<0x114, 0xf>(v));
  v = umax64(v, dpp_u64<0x118, 0xf>(v));
  v = umax64(v, dpp_u64<0x142, 0xa>(v));
  v = umax64(v, dpp_u64<0x143, 0xc>(v));
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v & 0xffffffffull), 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((u64)hi << 32) | lo;
}

// ---- single-workgroup leaf (h <= 1024) ---------------------------------------
// Thread t owns the R consecutive panel rows t*R + q, W register slots each,
// so lane order is row order.  Per column: the wave maxima go to LDS, one
// barrier, every thread folds them (rows ascend with the wave index, so a
// strict '>' keeps the lowest row) and the winning lane copies its row to LDS
// (LDS instruction count, not bandwidth, bounds this loop); then interchange,
// scaling, update, and a left rotation of each register row: column k+1
// moves to slot 0 and the finished column parks at the top, so after W
// rotations slot c holds column c and each row is stored as one contiguous
// run.
template <class T, int W, int R, int NTH>
__global__ __launch_bounds__(NTH) void getf2_panel_kernel(T* __restrict__ Tm, long ldt, int j0, int w, int m,
                                                         int* __restrict__ ipiv, int* __restrict__ info) {
  constexpr int NW = NTH / 64;
  __shared__ u64 s_wkey[NW];
  __shared__ int s_wrow[NW];
  __shared__ __align__(16) T s_wcand[1][W];
  __shared__ __align__(16) T s_rowk[W];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int h = m - j0;
  const int kmax = min(h, w);
  const T sfmin = Sfmin<T>::v;

  T v[R][W];
  int lr[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    lr[q] = tid * R + q;
    const T* col = Tm + (long)(j0 + lr[q]) * ldt + j0;
    if (lr[q] < h) {
#pragma unroll
      for (int c = 0; c < W; ++c) v[q][c] = c < w ? col[c] : T(0);
    } else {
#pragma unroll
      for (int c = 0; c < W; ++c) v[q][c] = T(0);
    }
  }

  for (int k = 0; k < kmax; ++k) {
    // 1. lane best over its R rows (lower q wins ties), then the wave's max
    //    key and the lowest lane holding it.
    u64 lk = 0;
    int lq = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool act = lr[q] >= k && lr[q] < h;
      const u64 key = act ? pivot_key(v[q][0], lr[q] == k) : 0ull;
      if (key > lk) { lk = key; lq = q; }
    }
    const u64 wmax = wave_max_u64(lk);
    const int wl = (int)__builtin_ctzll(__ballot(lk == wmax));
    if (lane == wl) {
      s_wkey[wave] = wmax;
      s_wrow[wave] = lr[lq];
    }
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (lr[q] == k) {
#pragma unroll
        for (int c = 0; c < W; ++c) s_rowk[c] = v[q][c];
      }
    __syncthreads();
    // 2. fold the waves (rows ascend with the wave index, so a strict '>'
    //    keeps the lowest row on ties); only the workgroup's winning lane
    //    writes its row to LDS.
    u64 bk = s_wkey[0];
    int bw = 0;
#pragma unroll
    for (int x = 1; x < NW; ++x) {
      const u64 kx = s_wkey[x];
      if (kx > bk) { bk = kx; bw = x; }
    }
    const int wrow = s_wrow[bw];
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (lr[q] == wrow) {
#pragma unroll
        for (int c = 0; c < W; ++c) s_wcand[0][c] = v[q][c];
      }
    __syncthreads();
    const int p = wrow;
    const T* prow_t = s_wcand[0];

    // 3. interchange, scale, rank-1 update
    const T piv = prow_t[0];
    if (tid == 0) {
      ipiv[j0 + k] = j0 + p + 1;
      if (piv == T(0)) atomicCAS(info, 0, j0 + k + 1);
    }
    const bool use_recip = fabs(piv) >= sfmin;
    const T rp = T(1) / piv;
    const int lim = W - k;  // slots 1..lim-1 hold unfinished columns
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (lr[q] == k) {
#pragma unroll
        for (int c = 0; c < W; ++c) v[q][c] = prow_t[c];
      } else if (lr[q] == p) {
#pragma unroll
        for (int c = 0; c < W; ++c) v[q][c] = s_rowk[c];
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (lr[q] > k && lr[q] < h) {
        T l = v[q][0];
        if (piv != T(0)) l = use_recip ? rp * l : l / piv;
        v[q][0] = l;
#pragma unroll
        for (int c = 1; c < W; ++c) {
          const T u = v[q][c] - l * prow_t[c];
          v[q][c] = (c < lim) ? u : v[q][c];
        }
      }
    }
    // 4. rotate
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const T t0 = v[q][0];
#pragma unroll
      for (int c = 0; c < W - 1; ++c) v[q][c] = v[q][c + 1];
      v[q][W - 1] = t0;
    }
    // Next column rewrites the LDS slots only after this barrier.
    __syncthreads();
  }
  for (int k = kmax; k < W; ++k) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const T t0 = v[q][0];
#pragma unroll
      for (int c = 0; c < W - 1; ++c) v[q][c] = v[q][c + 1];
      v[q][W - 1] = t0;
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (lr[q] >= h) continue;
    T* col = Tm + (long)(j0 + lr[q]) * ldt + j0;
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < w) col[c] = v[q][c];
  }
}

// ---- the exchange words of the multi-workgroup leaf --------------------------
// Every 64-bit word of a slot carries 32 bits of payload and the 32-bit flag
// tagbase + k + 1 of the column it belongs to (a double travels as two
// words), so a reader validates each word it loads and needs no fence or tag
// from the writer -- 64-bit stores are single-copy atomic.  A workgroup
// cannot publish column k+1 before every workgroup has read column k, so two
// buffers (column parity) suffice.  The exchange buffer must be zero before
// the first launch of a factorization (flags only grow after that: tagbase
// advances by 64 per launch).
constexpr int LL_SMAX = 512;                 // slots (GMAX workgroups, room for 8 per workgroup)
constexpr int LL_SLOT = 4 + 2 * WMAX;        // words: key lo, key hi, row, pad, values (2 words each)
constexpr size_t LL_PAR = (size_t)(LL_SMAX + 1) * LL_SLOT;  // slots + the row-k slot

__device__ __forceinline__ u64 ll_word(unsigned payload, unsigned flag) {
  return ((u64)flag << 32) | (u64)payload;
}
__device__ __forceinline__ void ll_put_double(u64* p, double v, unsigned flag) {
  const u64 b = (u64)__double_as_longlong(v);
  st_dev(p, ll_word((unsigned)(b & 0xffffffffull), flag));
  st_dev(p + 1, ll_word((unsigned)(b >> 32), flag));
}
// true when both words carry the flag; *v receives the double
__device__ __forceinline__ bool ll_get_double(const u64* p, unsigned flag, double* v) {
  const u64 lo = ld_dev(p), hi = ld_dev(p + 1);
  *v = __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
  return (unsigned)(lo >> 32) == flag && (unsigned)(hi >> 32) == flag;
}

// ---- multi-workgroup leaf: flag-in-data exchange, column loop unrolled by 8 ---
//
// The 64 columns run as 8 groups of 8: inside a group the column's register
// slot is a compile-time constant kk (no per-column rotation or mask: the
// update touches slots kk+1..7 of chunk 0 and the whole 8-slot chunks
// 1..7-g that still hold unfinished columns, selected by a uniform branch per
// chunk); after each group every row moves by 8 slots at once (measured: ~2 us
// per column less vector work than a one-slot rotation with masks).  One
// candidate per workgroup is exchanged:
//   B1  every wave leaves (key, row) in LDS; all threads fold the eight;
//       the winning wave stages its winning lane's row in LDS and stores
//       it, flag-tagged, to the workgroup's slot (no barrier, no wait);
//   the wave holding row k publishes it after the candidate;
//   wave 0 polls the G slot headers, takes the max key / lowest slot, loads
//       the winner's row and row k (flag-validated), hands them over
//       through LDS;
//   B2  all threads interchange, scale and update.
// LDS hand-over buffers alternate with the column parity, so a wave may run
// one column ahead without a barrier at the end of each column.
template <class T, int NTH, bool XCD0>
__global__ __launch_bounds__(NTH) void getf2_panel_w8_kernel(T* __restrict__ Tm, long ldt, int j0, int w, int m,
                                                            int* __restrict__ ipiv, int* __restrict__ info,
                                                            unsigned tagbase, u64* __restrict__ xbuf,
                                                            int* __restrict__ err, int Gp,
                                                            unsigned* __restrict__ ctr) {
  constexpr int W = 64;
  constexpr int NW = NTH / 64;
  __shared__ __align__(16) T s_rowk[2][W];
  __shared__ __align__(16) T s_prow[2][W];
  __shared__ __align__(16) T s_pub[NW][W];
  __shared__ u64 s_wkey[NW];
  __shared__ int s_wrow[NW];
  __shared__ int s_p[2];
  __shared__ int s_rank;
  const int tid = threadIdx.x;
  int G = gridDim.x;
  int b = blockIdx.x;
  if (XCD0) {
    // The blocks that run on XCD 0 (hardware register XCC_ID) take ranks in
    // arrival order (ctr[0]); the first Gp of them take part, rank r owning
    // rows [r*NTH, r*NTH + NTH), so the panel owns XCD 0's CUs while the
    // update leaves them alone; every other block leaves.  Each block then
    // counts its arrival (ctr[1]); the last to arrive -- every rank decision
    // is then made -- sets *err when fewer than Gp blocks ran on XCD 0, so
    // that every wait gives up and the call reports the failure instead of
    // leaving rows unfactored, and resets both counters for the next launch.
    if (tid == 0) {
      int r = -1;
      if ((__builtin_amdgcn_s_getreg(6164) & 15u) == 0u)  // HW_REG_XCC_ID
        r = (int)__hip_atomic_fetch_add(&ctr[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned arrived = __hip_atomic_fetch_add(&ctr[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (arrived == gridDim.x - 1) {
        const unsigned ranks = __hip_atomic_load(&ctr[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (ranks < (unsigned)Gp) st_dev(err, 1);
        __hip_atomic_store(&ctr[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctr[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      s_rank = (r >= 0 && r < Gp) ? r : -1;
    }
    __syncthreads();
    if (s_rank < 0) return;
    G = Gp;
    b = s_rank;
  }
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int h = m - j0;
  const int kmax = min(h, w);
  const int wrow0 = b * NTH + wave * 64;
  const T sfmin = Sfmin<T>::v;
  // Publishes the row held by lane src.  With a pending (deferred) update of
  // the previous column -- multiplier lp of the source row, pivot row pr --
  // the columns c with cmin < c < cmax are published as updated, so readers
  // get the row as the reference has it at this column.
  auto publish_row = [&](u64* dst, int src, unsigned fl, const T* vals, bool pend, T lp, bool ap, const T* pr,
                         int cmin, int cmax) {
    if (lane == src) {
#pragma unroll
      for (int c = 0; c < W; ++c) s_pub[wave][c] = vals[c];
    }
    __builtin_amdgcn_wave_barrier();
    T val = s_pub[wave][lane];
    if (pend) {
      const T ls = __shfl(lp, src);
      const bool as = __shfl((int)ap, src) != 0;
      if (as && lane > cmin && lane < cmax) val = val - ls * pr[lane];
    }
    ll_put_double(dst + 4 + 2 * lane, (double)val, fl);
    __builtin_amdgcn_wave_barrier();
  };

  T v[W];
  const int lr = b * NTH + tid;
  {
    const T* col = Tm + (long)(j0 + lr) * ldt + j0;
    if (lr < h) {
#pragma unroll
      for (int c = 0; c < W; ++c) v[c] = c < w ? col[c] : T(0);
    } else {
#pragma unroll
      for (int c = 0; c < W; ++c) v[c] = T(0);
    }
  }
  // the row's loads complete here, before the column loop: left pending,
  // they would make the compiler wait for every outstanding memory
  // operation (this column's exchange stores included) wherever the loop
  // first reads v, such as the candidate row's publication
  wait_vm();
  bool failed = false;
  int rotated = 0;
  // look-ahead: after column k's exchange a row updates only column k+1 (the
  // next pivot column) before the next pivot search; the rest of column k's
  // update is deferred until this workgroup has published column k+1's
  // candidate, so it runs while the exchange is in flight.  Not across the
  // end of a group (the rotation) nor after the leaf's last column.
  bool pend = false;   // a deferred update of the previous column is due (uniform)
  T l_prev = T(0);     // its multiplier for this thread's row
  bool act_prev = false;
  for (int g = 0; g < 8 && 8 * g < kmax; ++g) {
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const int k = 8 * g + kk;
      if (k < kmax) {
        const int par = k & 1;
        const unsigned flag = tagbase + (unsigned)k + 1u;
        u64* buf = xbuf + (size_t)par * LL_PAR;
        // 1. wave candidates -> LDS
        const bool act = lr >= k && lr < h;
        const u64 key = act ? pivot_key(v[kk], lr == k) : 0ull;
        const u64 wmax = wave_max_u64(key);
        const int wl = (int)__builtin_ctzll(__ballot(key == wmax));
        if (lane == wl) {
          s_wkey[wave] = wmax;
          s_wrow[wave] = lr;
        }
        __syncthreads();
        // 2. fold; the winning wave publishes the workgroup's candidate
        u64 bk = s_wkey[0];
        int bw = 0;
#pragma unroll
        for (int x = 1; x < NW; ++x) {
          const u64 kx = s_wkey[x];
          if (kx > bk) { bk = kx; bw = x; }
        }
        if (wave == bw) {
          // header first, then the row (readers check every word's flag)
          u64* sl = buf + (size_t)b * LL_SLOT;
          if (lane == 0) {
            st_dev(sl + 0, ll_word((unsigned)(bk & 0xffffffffull), flag));
            st_dev(sl + 1, ll_word((unsigned)(bk >> 32), flag));
            st_dev(sl + 2, ll_word((unsigned)s_wrow[bw], flag));
          }
          if (bk != 0ull) publish_row(sl, wl, flag, v, pend, l_prev, act_prev, s_prow[par ^ 1], kk, 8 * (8 - g));
        }
        // row k, for the interchange, after the candidate: only the owner of
        // the pivot row needs it, once the exchange is over, so it stays off
        // the path to the workgroup's header (measured: publishing it first
        // made the workgroup holding the panel's top rows the last to
        // publish in every column, 117.5 -> 121.5 ms at n = 16384)
        if (k >= wrow0 && k < wrow0 + 64)
          publish_row(buf + (size_t)LL_SMAX * LL_SLOT, k - wrow0, flag, v, pend, l_prev, act_prev, s_prow[par ^ 1], kk,
                      8 * (8 - g));
        // the previous column's deferred update, while the exchange is in
        // flight (its pivot row stays in s_prow[par ^ 1] until barrier 2)
        if (pend && act_prev) {
#pragma unroll
          for (int c = kk + 1; c < 8; ++c) v[c] = v[c] - l_prev * s_prow[par ^ 1][c];
#pragma unroll
          for (int q = 1; q < 8; ++q) {
            if (q < 8 - g) {
#pragma unroll
              for (int c = 8 * q; c < 8 * q + 8; ++c) v[c] = v[c] - l_prev * s_prow[par ^ 1][c];
            }
          }
        }
        // 3. wave 0: the G candidates, the winner's row and row k, in one
        // polling loop: with every poll it also loads row k and the row of
        // the best candidate seen so far (max key, lowest slot -- the final
        // rule once every header is in), so when the last header does not
        // change the winner, its row is already there
        if (wave == 0) {
          u64 kv = 0ull;
          bool kok = lane >= G;
          double rkv = 0.0, pv = 0.0;
          bool rk_ok = false, ok = false, rok = false;
          u64 roww = 0;
          int spec = -1;
          const u64* rks = buf + (size_t)LL_SMAX * LL_SLOT;
          unsigned spins = 0;
          while (!failed) {
            if (!kok) {
              const u64* sl = buf + (size_t)lane * LL_SLOT;
              const u64 lo = ld_dev(sl), hi = ld_dev(sl + 1);
              if ((unsigned)(lo >> 32) == flag && (unsigned)(hi >> 32) == flag) {
                kv = ((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull);
                kok = true;
              }
            }
            if (!rk_ok) rk_ok = ll_get_double(rks + 4 + 2 * lane, flag, &rkv);
            const u64 km = wave_max_u64(kv);
            int best = (int)__builtin_ctzll(__ballot(kv == km));
            if (best >= G) best = 0;
            if (best != spec) {
              spec = best;
              ok = false;
              rok = false;
            }
            const u64* src = buf + (size_t)spec * LL_SLOT;
            if (!ok) ok = ll_get_double(src + 4 + 2 * lane, flag, &pv);
            if (!rok) {
              roww = ld_dev(src + 2);
              rok = (unsigned)(roww >> 32) == flag;
            }
            if (__all(kok && ok && rok && rk_ok)) break;
            __builtin_amdgcn_s_sleep(1);
            ++spins;
            if ((spins & 255u) == 1u && ld_dev(err) != 0) failed = true;
            if (spins > (1u << 20)) { if (lane == 0) st_dev(err, 1); failed = true; }
          }
          int prow = (int)(roww & 0xffffffffull);
          if (prow < k || prow >= h) prow = k;  // only after a timeout
          s_prow[par][lane] = (T)pv;
          s_rowk[par][lane] = (T)rkv;
          if (lane == 0) s_p[par] = prow;
        }
        __syncthreads();
        const int p = s_p[par];
        // 4. interchange, scale, rank-1 update of the unfinished slots
        const T piv = s_prow[par][kk];
        if (b == 0 && tid == 0) {
          ipiv[j0 + k] = j0 + p + 1;
          if (piv == T(0)) atomicCAS(info, 0, j0 + k + 1);
        }
        if (lr == k) {
#pragma unroll
          for (int c = 0; c < W; ++c) v[c] = s_prow[par][c];
        } else if (lr == p) {
#pragma unroll
          for (int c = 0; c < W; ++c) v[c] = s_rowk[par][c];
        }
        // defer the rest of this column's update when column k+1 is in this
        // group and is not the leaf's last column's successor
        const bool defer = kk < 7 && k + 1 < kmax;
        act_prev = lr > k && lr < h;
        if (act_prev) {
          T l = v[kk];
          if (piv != T(0)) l = (fabs(piv) >= sfmin) ? (T(1) / piv) * l : l / piv;
          v[kk] = l;
          l_prev = l;
          if (defer) {
            if (kk < 7) v[kk + 1] = v[kk + 1] - l * s_prow[par][kk + 1];
          } else {
#pragma unroll
            for (int c = kk + 1; c < 8; ++c) v[c] = v[c] - l * s_prow[par][c];
#pragma unroll
            for (int q = 1; q < 8; ++q) {
              if (q < 8 - g) {
#pragma unroll
                for (int c = 8 * q; c < 8 * q + 8; ++c) v[c] = v[c] - l * s_prow[par][c];
              }
            }
          }
        }
        pend = defer;
      }
    }
    {
      T t[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) t[c] = v[c];
#pragma unroll
      for (int c = 0; c < W - 8; ++c) v[c] = v[c + 8];
#pragma unroll
      for (int c = 0; c < 8; ++c) v[W - 8 + c] = t[c];
    }
    ++rotated;
  }
  for (; rotated < 8; ++rotated) {
    T t[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) t[c] = v[c];
#pragma unroll
    for (int c = 0; c < W - 8; ++c) v[c] = v[c + 8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[W - 8 + c] = t[c];
  }
  if (lr < h) {
    T* col = Tm + (long)(j0 + lr) * ldt + j0;
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < w) col[c] = v[c];
  }
}

// ---- tall leaf: one column per two launches -----------------------------------
// For panels taller than GMAX register-resident workgroups hold, and for the
// leaves of a factorization redone after a bounded wait gave up: the leaf
// stays in HBM and no workgroup ever waits for another.  Column k:
//   tall_final   one workgroup folds the per-workgroup maxima of column k
//                (IDAMAX keys as above), writes IPIV and INFO, and copies the
//                pivot row and row k of the leaf to the scratch;
//   tall_update  every row r > k of the leaf: the interchange, the scaling and
//                the rank-1 update, in the same operations as the register
//                forms; row k takes the pivot row.  Each workgroup also
//                leaves the maximum key of column k+1 over its rows for the
//                next tall_final.
// tall_keys gives column 0 its maxima.  Rows past the leaf's w columns are
// never touched.
constexpr int TALL_NB = 512;  // workgroups of tall_update / tall_keys (at most)

struct TallBuf {
  u64 key[TALL_NB];
  int row[TALL_NB];
  double prow[WMAX];
  double rowk[WMAX];
  int p;
};

__device__ __forceinline__ void key_fold(u64& k, int& r, u64 k2, int r2) {
  if (k2 > k || (k2 == k && r2 < r)) { k = k2; r = r2; }
}

// workgroup maximum of (key, row) -> tb->key/row[blockIdx.x]
__device__ __forceinline__ void tall_block_max(u64 k, int r, TallBuf* tb) {
  __shared__ u64 sk[4];
  __shared__ int sr[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const u64 k2 = ((u64)(unsigned)__shfl_xor((int)(unsigned)(k >> 32), off) << 32) |
                   (u64)(unsigned)__shfl_xor((int)(unsigned)(k & 0xffffffffull), off);
    const int r2 = __shfl_xor(r, off);
    key_fold(k, r, k2, r2);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sk[wave] = k; sr[wave] = r; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int x = 1; x < (int)(blockDim.x >> 6); ++x) key_fold(k, r, sk[x], sr[x]);
    tb->key[blockIdx.x] = k;
    tb->row[blockIdx.x] = r;
  }
}

template <class T>
__global__ __launch_bounds__(256) void tall_keys_kernel(const T* __restrict__ Tm, long ldt, int j0, int k, int h,
                                                        TallBuf* __restrict__ tb) {
  u64 best = 0ull;
  int brow = 0x7fffffff;
  for (long r = k + (long)blockIdx.x * 256 + threadIdx.x; r < h; r += (long)gridDim.x * 256) {
    const u64 key = pivot_key(Tm[(long)(j0 + r) * ldt + j0 + k], r == k);
    key_fold(best, brow, key, (int)r);
  }
  tall_block_max(best, brow, tb);
}

template <class T>
__global__ __launch_bounds__(256) void tall_final_kernel(const T* __restrict__ Tm, long ldt, int j0, int k, int w,
                                                         int nparts, int* __restrict__ ipiv, int* __restrict__ info,
                                                         TallBuf* __restrict__ tb) {
  __shared__ u64 sk[4];
  __shared__ int sr[4];
  __shared__ int s_p;
  u64 best = 0ull;
  int brow = 0x7fffffff;
  for (int x = threadIdx.x; x < nparts; x += 256) key_fold(best, brow, tb->key[x], tb->row[x]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const u64 k2 = ((u64)(unsigned)__shfl_xor((int)(unsigned)(best >> 32), off) << 32) |
                   (u64)(unsigned)__shfl_xor((int)(unsigned)(best & 0xffffffffull), off);
    const int r2 = __shfl_xor(brow, off);
    key_fold(best, brow, k2, r2);
  }
  if ((threadIdx.x & 63) == 0) { sk[threadIdx.x >> 6] = best; sr[threadIdx.x >> 6] = brow; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int x = 1; x < 4; ++x) key_fold(best, brow, sk[x], sr[x]);
    // every active row has a key >= 1, so the fold always names a row of [k, h)
    const int p = (brow >= k && brow != 0x7fffffff) ? brow : k;
    s_p = p;
    tb->p = p;
    ipiv[j0 + k] = j0 + p + 1;
  }
  __syncthreads();
  const int p = s_p;
  if ((int)threadIdx.x < w) {
    const T pv = Tm[(long)(j0 + p) * ldt + j0 + threadIdx.x];
    tb->prow[threadIdx.x] = (double)pv;
    tb->rowk[threadIdx.x] = (double)Tm[(long)(j0 + k) * ldt + j0 + threadIdx.x];
    if ((int)threadIdx.x == k && pv == T(0)) atomicCAS(info, 0, j0 + k + 1);
  }
}

template <class T>
__global__ __launch_bounds__(256) void tall_update_kernel(T* __restrict__ Tm, long ldt, int j0, int k, int w, int h,
                                                          TallBuf* __restrict__ tb, int next_keys) {
  __shared__ T s_prow[WMAX], s_rowk[WMAX];
  const int p = tb->p;
  if ((int)threadIdx.x < WMAX) {
    s_prow[threadIdx.x] = (int)threadIdx.x < w ? (T)tb->prow[threadIdx.x] : T(0);
    s_rowk[threadIdx.x] = (int)threadIdx.x < w ? (T)tb->rowk[threadIdx.x] : T(0);
  }
  __syncthreads();
  const T piv = s_prow[k];
  const T sfmin = Sfmin<T>::v;
  u64 best = 0ull;
  int brow = 0x7fffffff;
  for (long r = k + (long)blockIdx.x * 256 + threadIdx.x; r < h; r += (long)gridDim.x * 256) {
    T* row = Tm + (long)(j0 + r) * ldt + j0;
    if (r == k) {
      for (int c = 0; c < w; ++c) row[c] = s_prow[c];
      continue;
    }
    const bool isp = r == p;
    T l = isp ? s_rowk[k] : row[k];
    if (piv != T(0)) l = (fabs(piv) >= sfmin) ? (T(1) / piv) * l : l / piv;
    if (isp)
      for (int c = 0; c < k; ++c) row[c] = s_rowk[c];
    row[k] = l;
    for (int c = k + 1; c < w; ++c) {
      const T x = isp ? s_rowk[c] : row[c];
      const T u = x - l * s_prow[c];
      row[c] = u;
      if (c == k + 1 && next_keys) key_fold(best, brow, pivot_key(u, r == k + 1), (int)r);
    }
  }
  if (next_keys) tall_block_max(best, brow, tb);
}

}  // namespace

// Launch configurations (measured on MI355X, per column of a 64-wide leaf):
// one 512-thread workgroup, one row x 64 per thread, up to 512 rows; one
// 512-thread workgroup, two rows x 32, up to 1024 rows; beyond that G
// workgroups of 512 threads, one row x 64 each; past GMAX workgroups the
// tall form.
int panel_max_width(int h) {
  if (h <= 512) return 64;
  if (h <= 1024) return 32;
  return 64;
}
static size_t tall_offset() { return (sizeof(u64) * 2 * LL_PAR + 255) / 256 * 256; }
// the XCD-0 panel's rank and arrival counters (zeroed with the exchange
// buffer at the start of a factorization; each launch resets its own)
static size_t rank_offset() { return tall_offset() + (sizeof(TallBuf) + 255) / 256 * 256; }
size_t panel_xchg_bytes(int) { return rank_offset() + 256; }

// Workgroups the panel kernel uses at height h (0 = a single workgroup).
int panel_groups(int h) { return h <= 1024 ? 0 : (h + 511) / 512; }

// mode: PANEL_PLAIN (G workgroups), PANEL_XCD0 (8*G blocks, those of the
// first XCDs take part), PANEL_TALL (the form without waits).
template <class T>
static int launch_panel_t(T* Tm, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase,
                          double* xchg, int* err, int mode, hipStream_t s) {
  const int h = m - j0;
  if (h <= 0 || w <= 0) return 0;
  if (w > panel_max_width(h)) return LA_ERR_LIMIT;
  if (h <= 512) {
    getf2_panel_kernel<T, 64, 1, 512><<<1, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info);
    return 0;
  }
  if (h <= 1024) {
    getf2_panel_kernel<T, 32, 2, 512><<<1, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info);
    return 0;
  }
  const int G = (h + 511) / 512;
  if (G > GMAX || mode == PANEL_TALL) {
    TallBuf* tb = reinterpret_cast<TallBuf*>(reinterpret_cast<char*>(xchg) + tall_offset());
    const int kmax = h < w ? h : w;
    auto parts = [&](int k) {
      const long rows = (long)h - k;
      const long g = (rows + 255) / 256;
      return (int)(g < TALL_NB ? (g < 1 ? 1 : g) : TALL_NB);
    };
    tall_keys_kernel<T><<<parts(0), 256, 0, s>>>(Tm, ldt, j0, 0, h, tb);
    for (int k = 0; k < kmax; ++k) {
      tall_final_kernel<T><<<1, 256, 0, s>>>(Tm, ldt, j0, k, w, parts(k), ipiv, info, tb);
      // the update's grid is also the next column's number of maxima
      tall_update_kernel<T><<<parts(k + 1 < kmax ? k + 1 : k), 256, 0, s>>>(Tm, ldt, j0, k, w, h, tb,
                                                                            k + 1 < kmax ? 1 : 0);
    }
    return 0;
  }
  u64* xb = reinterpret_cast<u64*>(xchg);
  if (mode == PANEL_XCD0)
    getf2_panel_w8_kernel<T, 512, true><<<(8 / (int)panel_xcds(G)) * G, 512, 0, s>>>(
        Tm, ldt, j0, w, m, ipiv, info, tagbase, xb, err, G,
        reinterpret_cast<unsigned*>(reinterpret_cast<char*>(xchg) + rank_offset()));
  else
    getf2_panel_w8_kernel<T, 512, false><<<G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xb, err, G,
                                                           nullptr);
  return 0;
}

int launch_getf2_panel_d(double* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase,
                         double* xchg, int* err, int mode, hipStream_t s) {
  return launch_panel_t<double>(T, ldt, j0, w, m, ipiv, info, tagbase, xchg, err, mode, s);
}
int launch_getf2_panel_s(float* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase,
                         double* xchg, int* err, int mode, hipStream_t s) {
  return launch_panel_t<float>(T, ldt, j0, w, m, ipiv, info, tagbase, xchg, err, mode, s);
}

}  // namespace la
