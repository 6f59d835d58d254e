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

namespace la {
namespace {

constexpr int GMAX = 64;            // workgroups per launch
constexpr int WMAX = 64;            // widest leaf
constexpr int SLOT = 2 * WMAX + 2;  // key, row, values[W], rowk[W]

template <class T> struct Sfmin;
template <> struct Sfmin<double> { static constexpr double v = 0x1p-1022; };
template <> struct Sfmin<float> { static constexpr float v = 0x1p-126f; };

typedef unsigned long long u64;

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

// XCD0: the grid holds 8*G blocks; only blocks running on XCD 0 take part
// (ranked in arrival order through *rankctr), so the panel owns that XCD's
// CUs while a concurrent trailing update keeps to the other seven.
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
      if ((__builtin_amdgcn_s_getreg(6164) & 15u) == 0u)  // HW_REG_XCC_ID
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
size_t panel_xchg_bytes(int) { return sizeof(double) * 2 * GMAX * SLOT; }
size_t panel_tag_count() { return GMAX; }

// Workgroups the panel kernel uses at height h (0 = a single workgroup).
int panel_groups(int h) { return h <= 1024 ? 0 : (h + 511) / 512; }

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
    if (xcd0 && G <= 32)  // one 512-thread workgroup per CU of XCD 0
      getf2_panel_kernel<T, 64, 1, false, 512, true><<<8 * G, 512, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, tagbase,
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
