// panel.hip -- cooperative leaf factorization of a tall panel (the unblocked
// core of DGETRF2, SRC/dgetrf2.f:179-214 applied column by column).
//
// The panel is A(j0:m-1, j0:j0+w-1) of the matrix being factored, held in the
// transposed workspace T = A^T, so matrix row r of the panel is T column r and
// its w values are contiguous.  The kernel keeps the whole panel in
// registers: G workgroups of NTH threads, thread t of workgroup b owning the
// R consecutive panel rows (b*NTH + t)*R + q, W slots each, so lane order is
// row order.  Per column k:
//   1. pivot search under the IDAMAX rule of BLAS/SRC/idamax.f:101-107 (first
//      index of max |x|, strict '>', a NaN at the head of the column wins, NaNs
//      elsewhere never do).  Each candidate gets a 64-bit key that orders
//      exactly like that rule: bits(|x|)+1 (IEEE bit patterns of non-negative
//      doubles sort like their values), row k's NaN -> bits(+inf)+1, other
//      NaNs -> 1, rows outside the active range -> 0; ties go to the lowest
//      row.  A wave takes the max key with DPP row shifts/broadcasts and the
//      lowest lane holding it with a ballot, and leaves (key, row) in LDS; the
//      owner of row k copies row k to LDS.
//   2. one barrier; every thread folds the wave results and the workgroup's
//      winning lane copies its row to LDS (LDS instruction count, not
//      bandwidth, bounds this loop, so only one lane per workgroup writes a
//      row).  G == 1: that is the pivot.  G > 1: wave 0 publishes {key, row,
//      values[W], row k (owner only)} with device-coherent stores, waits for
//      them, raises its tag to tagbase+k+1, polls all G tags (bounded), takes
//      the max key over the G slots with the same DPP/ballot rule and loads
//      the winning slot's row; every workgroup reaches the same decision.
//   3. interchange (row k takes the pivot row, the pivot row takes the old row
//      k: SRC/dgetrf2.f:196-200), scaling by the reciprocal of the pivot, or
//      division when |pivot| < sfmin (SRC/dgetrf2.f:204-210), and the rank-1
//      update of the remaining columns (SRC/dgetrf2.f:245-246 one column at a
//      time), applied also for a zero pivot as the reference's DGEMM has no
//      zero test;
//   4. a left rotation of each register row: column k+1 moves to slot 0 and
//      the finished column parks at the top; after W rotations slot c holds
//      column c and each row is stored as one contiguous run.
// IPIV receives absolute 1-based rows; INFO the first exact zero pivot
// (SRC/dgetrf2.f:192,212-214).  Every wait is bounded: a timeout sets *err,
// later waits give up at once, and the kernel always runs to its end.
#include "la_common.h"

#include <cstdlib>
#include <cstring>

namespace la {
namespace {

constexpr int GMAX = 64;            // workgroups per launch
constexpr int WMAX = 64;            // widest leaf
constexpr int SLOT = 2 * WMAX + 2;  // key, row, values[W], rowk[W]

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

// XCD0: the grid holds (8 / panel_xcds(G)) * G blocks; only blocks running
// on the first panel_xcds(G) XCDs take part (ranked in arrival order through
// *rankctr), so the panel owns those XCDs' CUs (one 512-thread workgroup per
// CU, 32 per XCD) while a concurrent trailing update keeps to the others.
template <class T, int W, int R, bool SOLO, int NTH, bool XCD0>
__global__ __launch_bounds__(NTH) void getf2_panel_kernel(T* __restrict__ Tm, long ldt, int j0, int w, int m,
                                                         int* __restrict__ ipiv, int* __restrict__ info,
                                                         unsigned tagbase, double* __restrict__ xbuf,
                                                         unsigned* __restrict__ tags, int* __restrict__ err,
                                                         int Gp, unsigned* __restrict__ rankctr) {
  constexpr int NW = NTH / 64;
  __shared__ u64 s_wkey[NW];
  __shared__ int s_wrow[NW];
  __shared__ __align__(16) T s_wcand[1][W];
  __shared__ __align__(16) T s_rowk[W];
  __shared__ __align__(16) T s_prow[W];
  __shared__ int s_p;

  __shared__ int s_rank;
  const int tid = threadIdx.x;
  int G = gridDim.x;
  int b = blockIdx.x;
  if (XCD0) {
    if (tid == 0) {
      int r = -1;
      if ((__builtin_amdgcn_s_getreg(6164) & 15u) < panel_xcds(Gp))  // HW_REG_XCC_ID
        r = (int)__hip_atomic_fetch_add(rankctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  const T sfmin = Sfmin<T>::v;

  T v[R][W];
  int lr[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    lr[q] = (b * NTH + tid) * R + q;
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
    int p;
    const T* prow_t = nullptr;
    if (SOLO) {
      p = wrow;
      prow_t = s_wcand[0];
    } else {
      const int par = k & 1;
      const unsigned want = tagbase + (unsigned)k + 1u;
      const int owner = k / (NTH * R);
      if (wave == 0) {
        double* my = xbuf + ((size_t)par * GMAX + b) * SLOT;
        if (lane < W) {
          st_dev(&my[2 + lane], (double)s_wcand[0][lane]);
          if (b == owner) st_dev(&my[2 + WMAX + lane], (double)s_rowk[lane]);
        }
        if (lane == 0) {
          st_dev(&my[0], __longlong_as_double((long long)bk));
          st_dev(&my[1], (double)wrow);
        }
        wait_vm();
        if (lane == 0) st_dev(&tags[b], want);
        unsigned spins = 0;
        while (true) {
          const bool mine = lane >= G || ld_dev(&tags[lane]) >= want;
          if (__all(mine)) break;
          __builtin_amdgcn_s_sleep(1);
          ++spins;
          if ((spins & 255u) == 0 && ld_dev(err) != 0) break;
          if (spins > (1u << 20)) { if (lane == 0) st_dev(err, 1); break; }
        }
        // keys of all workgroups -> winner (workgroups hold ascending rows:
        // the lowest lane holding the max key is the lowest row)
        const u64 gk = lane < G ? (u64)__double_as_longlong(ld_dev(&xbuf[((size_t)par * GMAX + lane) * SLOT])) : 0ull;
        const u64 gmax = wave_max_u64(gk);
        const int gg = (int)__builtin_ctzll(__ballot(gk == gmax));
        const double* src = xbuf + ((size_t)par * GMAX + gg) * SLOT;
        const double prw = ld_dev(&src[1]);
        double pv = 0.0, rk = 0.0;
        if (lane < W) {
          pv = ld_dev(&src[2 + lane]);
          if (b != owner) rk = ld_dev(&xbuf[((size_t)par * GMAX + owner) * SLOT + 2 + WMAX + lane]);
        }
        if (lane < W) {
          s_prow[lane] = (T)pv;
          if (b != owner) s_rowk[lane] = (T)rk;
        }
        if (lane == 0) s_p = (int)prw;
      }
      __syncthreads();
      p = s_p;
      prow_t = s_prow;
    }

    // 3. interchange, scale, rank-1 update
    const T piv = prow_t[0];
    if (b == 0 && tid == 0) {
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


// ---- multi-workgroup leaf with a flag-in-data exchange -----------------------
//
// Same arithmetic as getf2_panel_kernel (same register layout, keys, swap,
// scale, update, rotation), but the cross-workgroup step is rebuilt around
// one exchange per column with no separate tag word:
//   * every wave publishes its own best candidate (no workgroup fold first):
//     the lane holding the wave's max key writes its key and, when the key is
//     non-zero, its row of W values to the wave's slot;
//   * every 64-bit word of a slot carries 32 bits of payload and the 32-bit
//     flag tagbase + k + 1 of the column it belongs to (a double travels as two
//     words), so a reader validates each word it loads and needs no fence or
//     tag from the writer -- 64-bit stores are single-copy atomic;
//   * wave 0 of each workgroup loads all slots' keys together with row k
//     (published by its owner at the end of the previous column), takes the
//     max key and the lowest slot holding it (slots ascend with the rows, so
//     that is the reference's lowest row), loads the winner's row and row
//     index, and hands them to its workgroup through LDS: one barrier per
//     column.  A workgroup cannot publish column k+1 before it has read all of
//     column k, so two buffers (column parity) suffice.
// The exchange buffer must be zero before the first launch of a
// factorization (flags only grow after that: tagbase advances by 64 per
// launch).  Every wait is bounded; a timeout sets *err and the kernel runs
// to its end with whatever values it has (indexes stay in range).
constexpr int LL_SMAX = 512;                 // wave slots: 64 workgroups x 8 waves
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

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

template <class T, int W, int NTH, bool XCD0>
__global__ __launch_bounds__(NTH) void getf2_panel_ll_kernel(T* __restrict__ Tm, long ldt, int j0, int w, int m,
                                                            int* __restrict__ ipiv, int* __restrict__ info,
                                                            unsigned tagbase, u64* __restrict__ xbuf,
                                                            int* __restrict__ err, int Gp,
                                                            unsigned* __restrict__ rankctr) {
  constexpr int NW = NTH / 64;
  __shared__ __align__(16) T s_rowk[W];
  __shared__ __align__(16) T s_prow[W];
  __shared__ __align__(16) T s_pub[NW][W];  // per-wave staging of a row being published
  __shared__ int s_p;
  __shared__ int s_rank;
  const int tid = threadIdx.x;
  int G = gridDim.x;
  int b = blockIdx.x;
  if (XCD0) {
    if (tid == 0) {
      int r = -1;
      if ((__builtin_amdgcn_s_getreg(6164) & 15u) < panel_xcds(Gp))  // HW_REG_XCC_ID
        r = (int)__hip_atomic_fetch_add(rankctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  const int S = G * NW;  // slots in use
  const int myslot = b * NW + wave;
  const int wrow0 = b * NTH + wave * 64;  // first row of this wave
  const T sfmin = Sfmin<T>::v;
  // A row held by lane `src` of this wave goes to global memory as one
  // coalesced store (lane c writes value c): the lane stages it in LDS, the
  // wave reads it back distributed (same wave: LDS accesses stay in order).
  auto publish_row = [&](u64* dst, int src, unsigned fl, const T* vals) {
    if (lane == src) {
#pragma unroll
      for (int c = 0; c < W; ++c) s_pub[wave][c] = vals[c];
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < W) ll_put_double(dst + 4 + 2 * lane, (double)s_pub[wave][lane], fl);
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
  bool failed = false;
  // row 0 of the leaf, for column 0's interchange
  if (wrow0 == 0) publish_row(xbuf + (size_t)LL_SMAX * LL_SLOT, 0, tagbase + 1u, v);

  for (int k = 0; k < kmax; ++k) {
    const unsigned flag = tagbase + (unsigned)k + 1u;
    u64* buf = xbuf + (size_t)(k & 1) * LL_PAR;
    // 1. the wave's best candidate, published by the lane holding it
    const bool act = lr >= k && lr < h;
    const u64 key = act ? pivot_key(v[0], lr == k) : 0ull;
    const u64 wmax = wave_max_u64(key);
    const int wl = (int)__builtin_ctzll(__ballot(key == wmax));
    {
      u64* sl = buf + (size_t)myslot * LL_SLOT;
      if (wmax != 0ull) {
        publish_row(sl, wl, flag, v);
        if (lane == wl) st_dev(sl + 2, ll_word((unsigned)lr, flag));
      }
      if (lane == 0) {
        st_dev(sl + 0, ll_word((unsigned)(wmax & 0xffffffffull), flag));
        st_dev(sl + 1, ll_word((unsigned)(wmax >> 32), flag));
      }
    }
    // 2. wave 0: all keys + row k, then the winner's row
    if (wave == 0) {
      constexpr int NSL = LL_SMAX / 64;
      u64 kv[NSL];
      unsigned okmask = 0;
      double rkv = 0.0;
      bool rk_ok = lane >= W;
      const u64* rks = buf + (size_t)LL_SMAX * LL_SLOT;
      unsigned spins = 0;
      while (!failed) {
#pragma unroll
        for (int j = 0; j < NSL; ++j) {
          const int sidx = lane + 64 * j;
          if (sidx < S && !(okmask & (1u << j))) {
            const u64* sl = buf + (size_t)sidx * LL_SLOT;
            const u64 lo = ld_dev(sl), hi = ld_dev(sl + 1);
            if ((unsigned)(lo >> 32) == flag && (unsigned)(hi >> 32) == flag) {
              kv[j] = ((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull);
              okmask |= 1u << j;
            }
          } else if (sidx >= S) {
            kv[j] = 0ull;
            okmask |= 1u << j;
          }
        }
        if (!rk_ok) rk_ok = ll_get_double(rks + 4 + 2 * lane, flag, &rkv);
        if (__all(okmask == (1u << NSL) - 1u && rk_ok)) break;
        __builtin_amdgcn_s_sleep(1);
        ++spins;
        if ((spins & 255u) == 0 && ld_dev(err) != 0) failed = true;
        if (spins > (1u << 20)) { if (lane == 0) st_dev(err, 1); failed = true; }
      }
      u64 bk = 0ull;
      int bs = 0x7fffffff;
#pragma unroll
      for (int j = 0; j < NSL; ++j) {
        const u64 kk = (okmask & (1u << j)) ? kv[j] : 0ull;
        if (kk > bk) { bk = kk; bs = lane + 64 * j; }
      }
      const u64 gmax = wave_max_u64(bk);
      int ws = wave_min_i32(bk == gmax ? bs : 0x7fffffff);
      if (ws < 0 || ws >= S) ws = 0;
      const u64* src = buf + (size_t)ws * LL_SLOT;
      double pv = 0.0;
      bool ok = lane >= W;
      u64 roww = 0;
      bool rok = false;
      spins = 0;
      while (!failed) {
        if (!ok) ok = ll_get_double(src + 4 + 2 * lane, flag, &pv);
        if (!rok) {
          roww = ld_dev(src + 2);
          rok = (unsigned)(roww >> 32) == flag;
        }
        if (__all(ok && rok)) break;
        __builtin_amdgcn_s_sleep(1);
        ++spins;
        if ((spins & 255u) == 0 && ld_dev(err) != 0) failed = true;
        if (spins > (1u << 20)) { if (lane == 0) st_dev(err, 1); failed = true; }
      }
      int prow = (int)(roww & 0xffffffffull);
      if (prow < k || prow >= h) prow = k;  // only after a timeout
      if (lane < W) {
        s_prow[lane] = (T)pv;
        s_rowk[lane] = (T)rkv;
      }
      if (lane == 0) s_p = prow;
    }
    __syncthreads();
    const int p = s_p;
    // 3. interchange, scale, rank-1 update (as getf2_panel_kernel)
    const T piv = s_prow[0];
    if (b == 0 && tid == 0) {
      ipiv[j0 + k] = j0 + p + 1;
      if (piv == T(0)) atomicCAS(info, 0, j0 + k + 1);
    }
    const bool use_recip = fabs(piv) >= sfmin;
    const T rp = T(1) / piv;
    const int lim = W - k;
    if (lr == k) {
#pragma unroll
      for (int c = 0; c < W; ++c) v[c] = s_prow[c];
    } else if (lr == p) {
#pragma unroll
      for (int c = 0; c < W; ++c) v[c] = s_rowk[c];
    }
    if (lr > k && lr < h) {
      T l = v[0];
      if (piv != T(0)) l = use_recip ? rp * l : l / piv;
      v[0] = l;
#pragma unroll
      for (int c = 1; c < W; ++c) {
        const T u = v[c] - l * s_prow[c];
        v[c] = (c < lim) ? u : v[c];
      }
    }
    // 4. rotate; the owner of row k+1 publishes it for the next column
    {
      const T t0 = v[0];
#pragma unroll
      for (int c = 0; c < W - 1; ++c) v[c] = v[c + 1];
      v[W - 1] = t0;
    }
    if (k + 1 < kmax && k + 1 >= wrow0 && k + 1 < wrow0 + 64)  // uniform over the wave
      publish_row(xbuf + (size_t)((k + 1) & 1) * LL_PAR + (size_t)LL_SMAX * LL_SLOT, k + 1 - wrow0, flag + 1u, v);
  }
  for (int k = kmax; k < W; ++k) {
    const T t0 = v[0];
#pragma unroll
    for (int c = 0; c < W - 1; ++c) v[c] = v[c + 1];
    v[W - 1] = t0;
  }
  if (lr < h) {
    T* col = Tm + (long)(j0 + lr) * ldt + j0;
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < w) col[c] = v[c];
  }
}


// ---- the same exchange with a column loop unrolled by eight ------------------
//
// getf2_panel_ll_kernel moves every register row by one slot per column and
// masks the finished slots in the update: about 250 of its ~450 vector
// instructions per column.  Here the 64 columns run as 8 groups of 8: inside
// a group the column's slot is a compile-time constant kk (no rotation, no
// mask: the update touches slots kk+1..7 of chunk 0 and the whole 8-slot
// chunks 1..7-g that still hold unfinished columns, selected by a uniform
// branch per chunk); after each group the row moves by 8 slots at once.
// Row k is published at the start of its column by the wave holding it.
template <class T, int NTH, bool XCD0>
__global__ __launch_bounds__(NTH) void getf2_panel_ll8_kernel(T* __restrict__ Tm, long ldt, int j0, int w, int m,
                                                             int* __restrict__ ipiv, int* __restrict__ info,
                                                             unsigned tagbase, u64* __restrict__ xbuf,
                                                             int* __restrict__ err, int Gp,
                                                             unsigned* __restrict__ rankctr) {
  constexpr int W = 64;
  constexpr int NW = NTH / 64;
  __shared__ __align__(16) T s_rowk[W];
  __shared__ __align__(16) T s_prow[W];
  __shared__ __align__(16) T s_pub[NW][W];
  __shared__ int s_p;
  __shared__ int s_rank;
  const int tid = threadIdx.x;
  int G = gridDim.x;
  int b = blockIdx.x;
  if (XCD0) {
    if (tid == 0) {
      int r = -1;
      if ((__builtin_amdgcn_s_getreg(6164) & 15u) < panel_xcds(Gp))  // HW_REG_XCC_ID
        r = (int)__hip_atomic_fetch_add(rankctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  const int S = G * NW;
  const int myslot = b * NW + wave;
  const int wrow0 = b * NTH + wave * 64;
  const T sfmin = Sfmin<T>::v;
  auto publish_row = [&](u64* dst, int src, unsigned fl, const T* vals) {
    if (lane == src) {
#pragma unroll
      for (int c = 0; c < W; ++c) s_pub[wave][c] = vals[c];
    }
    __builtin_amdgcn_wave_barrier();
    ll_put_double(dst + 4 + 2 * lane, (double)s_pub[wave][lane], fl);
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
  bool failed = false;
  int rotated = 0;
  for (int g = 0; g < 8 && 8 * g < kmax; ++g) {
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const int k = 8 * g + kk;
      if (k < kmax) {
        const unsigned flag = tagbase + (unsigned)k + 1u;
        u64* buf = xbuf + (size_t)(k & 1) * LL_PAR;
        // row k, for the interchange
        if (k >= wrow0 && k < wrow0 + 64) publish_row(buf + (size_t)LL_SMAX * LL_SLOT, k - wrow0, flag, v);
        // 1. the wave's best candidate
        const bool act = lr >= k && lr < h;
        const u64 key = act ? pivot_key(v[kk], lr == k) : 0ull;
        const u64 wmax = wave_max_u64(key);
        const int wl = (int)__builtin_ctzll(__ballot(key == wmax));
        {
          u64* sl = buf + (size_t)myslot * LL_SLOT;
          if (wmax != 0ull) {
            publish_row(sl, wl, flag, v);
            if (lane == wl) st_dev(sl + 2, ll_word((unsigned)lr, flag));
          }
          if (lane == 0) {
            st_dev(sl + 0, ll_word((unsigned)(wmax & 0xffffffffull), flag));
            st_dev(sl + 1, ll_word((unsigned)(wmax >> 32), flag));
          }
        }
        // 2. wave 0: all keys + row k, then the winner's row
        if (wave == 0) {
          constexpr int NSL = LL_SMAX / 64;
          u64 kv[NSL];
          unsigned okmask = 0;
          double rkv = 0.0;
          bool rk_ok = false;
          const u64* rks = buf + (size_t)LL_SMAX * LL_SLOT;
          unsigned spins = 0;
          while (!failed) {
#pragma unroll
            for (int j = 0; j < NSL; ++j) {
              const int sidx = lane + 64 * j;
              if (sidx < S && !(okmask & (1u << j))) {
                const u64* sl = buf + (size_t)sidx * LL_SLOT;
                const u64 lo = ld_dev(sl), hi = ld_dev(sl + 1);
                if ((unsigned)(lo >> 32) == flag && (unsigned)(hi >> 32) == flag) {
                  kv[j] = ((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull);
                  okmask |= 1u << j;
                }
              } else if (sidx >= S) {
                kv[j] = 0ull;
                okmask |= 1u << j;
              }
            }
            if (!rk_ok) rk_ok = ll_get_double(rks + 4 + 2 * lane, flag, &rkv);
            if (__all(okmask == (1u << NSL) - 1u && rk_ok)) break;
            __builtin_amdgcn_s_sleep(1);
            ++spins;
            if ((spins & 255u) == 0 && ld_dev(err) != 0) failed = true;
            if (spins > (1u << 20)) { if (lane == 0) st_dev(err, 1); failed = true; }
          }
          u64 bk = 0ull;
          int bs = 0x7fffffff;
#pragma unroll
          for (int j = 0; j < NSL; ++j) {
            const u64 kk2 = (okmask & (1u << j)) ? kv[j] : 0ull;
            if (kk2 > bk) { bk = kk2; bs = lane + 64 * j; }
          }
          const u64 gmax = wave_max_u64(bk);
          int ws = wave_min_i32(bk == gmax ? bs : 0x7fffffff);
          if (ws < 0 || ws >= S) ws = 0;
          const u64* src = buf + (size_t)ws * LL_SLOT;
          double pv = 0.0;
          bool ok = false;
          u64 roww = 0;
          bool rok = false;
          spins = 0;
          while (!failed) {
            if (!ok) ok = ll_get_double(src + 4 + 2 * lane, flag, &pv);
            if (!rok) {
              roww = ld_dev(src + 2);
              rok = (unsigned)(roww >> 32) == flag;
            }
            if (__all(ok && rok)) break;
            __builtin_amdgcn_s_sleep(1);
            ++spins;
            if ((spins & 255u) == 0 && ld_dev(err) != 0) failed = true;
            if (spins > (1u << 20)) { if (lane == 0) st_dev(err, 1); failed = true; }
          }
          int prow = (int)(roww & 0xffffffffull);
          if (prow < k || prow >= h) prow = k;  // only after a timeout
          s_prow[lane] = (T)pv;
          s_rowk[lane] = (T)rkv;
          if (lane == 0) s_p = prow;
        }
        __syncthreads();
        const int p = s_p;
        // 3. interchange, scale, rank-1 update of the unfinished slots
        const T piv = s_prow[kk];
        if (b == 0 && tid == 0) {
          ipiv[j0 + k] = j0 + p + 1;
          if (piv == T(0)) atomicCAS(info, 0, j0 + k + 1);
        }
        if (lr == k) {
#pragma unroll
          for (int c = 0; c < W; ++c) v[c] = s_prow[c];
        } else if (lr == p) {
#pragma unroll
          for (int c = 0; c < W; ++c) v[c] = s_rowk[c];
        }
        if (lr > k && lr < h) {
          T l = v[kk];
          if (piv != T(0)) l = (fabs(piv) >= sfmin) ? (T(1) / piv) * l : l / piv;
          v[kk] = l;
#pragma unroll
          for (int c = kk + 1; c < 8; ++c) v[c] = v[c] - l * s_prow[c];
#pragma unroll
          for (int q = 1; q < 8; ++q) {
            if (q < 8 - g) {
#pragma unroll
              for (int c = 8 * q; c < 8 * q + 8; ++c) v[c] = v[c] - l * s_prow[c];
            }
          }
        }
      }
    }
    // move the row by 8 slots: the next group's columns to slots 0..7
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


// ---- workgroup-level flag-in-data exchange, column loop unrolled by eight ---
//
// Measured per column, the unrolled loop of getf2_panel_ll8_kernel saves
// about 2 us of vector work over the one-slot rotation, while its per-wave
// slots (8 per workgroup) cost exchange traffic.  This kernel keeps the
// unrolled loop and exchanges one candidate per workgroup:
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
                                                            unsigned* __restrict__ rankctr) {
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
    if (tid == 0) {
      int r = -1;
      if ((__builtin_amdgcn_s_getreg(6164) & 15u) < panel_xcds(Gp))  // HW_REG_XCC_ID
        r = (int)__hip_atomic_fetch_add(rankctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  auto publish_row = [&](u64* dst, int src, unsigned fl, const T* vals) {
    if (lane == src) {
#pragma unroll
      for (int c = 0; c < W; ++c) s_pub[wave][c] = vals[c];
    }
    __builtin_amdgcn_wave_barrier();
    ll_put_double(dst + 4 + 2 * lane, (double)s_pub[wave][lane], fl);
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
  bool failed = false;
  int rotated = 0;
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
          // header first: the readers' first poll needs only the keys, the
          // row is fetched by their second
          u64* sl = buf + (size_t)b * LL_SLOT;
          if (lane == 0) {
            st_dev(sl + 0, ll_word((unsigned)(bk & 0xffffffffull), flag));
            st_dev(sl + 1, ll_word((unsigned)(bk >> 32), flag));
            st_dev(sl + 2, ll_word((unsigned)s_wrow[bw], flag));
          }
          if (bk != 0ull) publish_row(sl, wl, flag, v);
        }
        // row k, for the interchange, after the candidate: only the owner of
        // the pivot row needs it, once the exchange is over, so it stays off
        // the path to the workgroup's header (measured: publishing it first
        // made the workgroup holding the panel's top rows the last to
        // publish in every column)
        if (k >= wrow0 && k < wrow0 + 64) publish_row(buf + (size_t)LL_SMAX * LL_SLOT, k - wrow0, flag, v);
        // 3. wave 0: the G candidates, then the winner's row and row k
        if (wave == 0) {
          u64 kv = 0ull;
          bool kok = lane >= G;
          double rkv = 0.0;
          bool rk_ok = false;
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
            if (__all(kok)) break;
            __builtin_amdgcn_s_sleep(1);
            ++spins;
            if ((spins & 255u) == 0 && ld_dev(err) != 0) failed = true;
            if (spins > (1u << 20)) { if (lane == 0) st_dev(err, 1); failed = true; }
          }
          const u64 gmax = wave_max_u64(kv);
          int gs = (int)__builtin_ctzll(__ballot(kv == gmax));
          if (gs >= G) gs = 0;
          const u64* src = buf + (size_t)gs * LL_SLOT;
          double pv = 0.0;
          bool ok = false;
          u64 roww = 0;
          bool rok = false;
          spins = 0;
          while (!failed) {
            if (!ok) ok = ll_get_double(src + 4 + 2 * lane, flag, &pv);
            if (!rok) {
              roww = ld_dev(src + 2);
              rok = (unsigned)(roww >> 32) == flag;
            }
            if (!rk_ok) rk_ok = ll_get_double(rks + 4 + 2 * lane, flag, &rkv);
            if (__all(ok && rok && rk_ok)) break;
            __builtin_amdgcn_s_sleep(1);
            ++spins;
            if ((spins & 255u) == 0 && ld_dev(err) != 0) failed = true;
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
        if (lr > k && lr < h) {
          T l = v[kk];
          if (piv != T(0)) l = (fabs(piv) >= sfmin) ? (T(1) / piv) * l : l / piv;
          v[kk] = l;
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


}  // namespace

// Launch configurations (measured on MI355X, per column of a 64-wide leaf):
// one 512-thread workgroup, one row x 64 per thread, up to 512 rows; one
// 512-thread workgroup, two rows x 32, up to 1024 rows; beyond that G
// workgroups of 512 threads, one row x 64 each.
int panel_max_width(int h) {
  if (h <= 512) return 64;
  if (h <= 1024) return 32;
  return 64;
}
size_t panel_xchg_bytes(int) {
  const size_t tagged = sizeof(double) * 2 * GMAX * SLOT;
  const size_t ll = sizeof(u64) * 2 * LL_PAR;
  return tagged > ll ? tagged : ll;
}
size_t panel_tag_count() { return GMAX; }

// Workgroups the panel kernel uses at height h (0 = a single workgroup).
int panel_groups(int h) { return h <= 1024 ? 0 : (h + 511) / 512; }

// Multi-workgroup panel variant: getf2_panel_w8_kernel by default;
// LAPACK_AMD_PANEL=ll8 / ll / tag select the per-wave-slot unrolled kernel, the per-wave-slot rotating kernel, or the tag-word exchange of
// getf2_panel_kernel (kept for comparison).
static int panel_variant() {
  static int v = -1;
  if (v < 0) {
    const char* e = std::getenv("LAPACK_AMD_PANEL");
    v = 3;
    if (e && std::strcmp(e, "tag") == 0) v = 0;
    if (e && std::strcmp(e, "ll") == 0) v = 1;
    if (e && std::strcmp(e, "ll8") == 0) v = 2;
  }
  return v;
}

template <class T>
static int launch_panel_t(T* Tm, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase,
                          double* xchg, unsigned* tags, int* err, unsigned* rankctr, bool xcd0, hipStream_t s) {
  const int h = m - j0;
  if (h <= 0 || w <= 0) return 0;
  if (w > panel_max_width(h)) return LA_ERR_LIMIT;
  if (h <= 512) {
    getf2_panel_kernel<T, 64, 1, true, 512, false><<<1, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xchg,
                                                                    tags, err, 1, rankctr);
  } else if (h <= 1024) {
    getf2_panel_kernel<T, 32, 2, true, 512, false><<<1, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xchg,
                                                                    tags, err, 1, rankctr);
  } else {
    const int G = (h + 511) / 512;
    if (G > GMAX) return LA_ERR_LIMIT;
    if (panel_variant() == 3) {
      u64* xb = reinterpret_cast<u64*>(xchg);
      if (xcd0 && G <= 64)
        getf2_panel_w8_kernel<T, 512, true><<<(8 / (int)panel_xcds(G)) * G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xb, err,
                                                                 G, rankctr);
      else
        getf2_panel_w8_kernel<T, 512, false><<<G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xb, err, G,
                                                              rankctr);
      return 0;
    }
    if (panel_variant() == 2) {
      u64* xb = reinterpret_cast<u64*>(xchg);
      if (xcd0 && G <= 64)
        getf2_panel_ll8_kernel<T, 512, true><<<(8 / (int)panel_xcds(G)) * G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xb, err,
                                                                  G, rankctr);
      else
        getf2_panel_ll8_kernel<T, 512, false><<<G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xb, err, G,
                                                               rankctr);
      return 0;
    }
    if (panel_variant() == 1) {
      u64* xb = reinterpret_cast<u64*>(xchg);
      if (xcd0 && G <= 64)
        getf2_panel_ll_kernel<T, 64, 512, true><<<(8 / (int)panel_xcds(G)) * G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xb,
                                                                     err, G, rankctr);
      else
        getf2_panel_ll_kernel<T, 64, 512, false><<<G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase, xb, err,
                                                                  G, rankctr);
      return 0;
    }
    if (xcd0 && G <= 64)  // one 512-thread workgroup per CU of XCD 0
      getf2_panel_kernel<T, 64, 1, false, 512, true><<<(8 / (int)panel_xcds(G)) * G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase,
                                                                          xchg, tags, err, G, rankctr);
    else
      getf2_panel_kernel<T, 64, 1, false, 512, false><<<G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase,
                                                                       xchg, tags, err, G, rankctr);
  }
  return 0;
}

int launch_getf2_panel_d(double* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase,
                         double* xchg, unsigned* tags, int* err, unsigned* rankctr, bool xcd0, hipStream_t s) {
  return launch_panel_t<double>(T, ldt, j0, w, m, ipiv, info, tagbase, xchg, tags, err, rankctr, xcd0, s);
}
int launch_getf2_panel_s(float* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase,
                         double* xchg, unsigned* tags, int* err, unsigned* rankctr, bool xcd0, hipStream_t s) {
  return launch_panel_t<float>(T, ldt, j0, w, m, ipiv, info, tagbase, xchg, tags, err, rankctr, xcd0, s);
}

}  // namespace la
