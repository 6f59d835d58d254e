// panel.hip -- cooperative leaf factorization of a tall panel (the unblocked
// core of DGETRF2, SRC/dgetrf2.f:179-214 applied column by column).
//
// The panel is A(j0:m-1, j0:j0+w-1) of the matrix being factored, held in the
// transposed workspace T = A^T, so matrix row r of the panel is T column r and
// each panel column is a contiguous run of T.  The kernel keeps the whole
// panel in registers: G workgroups of 256 threads, each thread owning R
// panel rows (w <= W values each).  Per column k it
//   1. finds its workgroup's candidate pivot: the IDAMAX rule of
//      BLAS/SRC/idamax.f:101-107 (first index of max |x|, strict '>', a NaN
//      at the head of the column wins, NaNs elsewhere never do) expressed as
//      a (key desc, row asc) reduction, first across the wave with DPP
//      shuffles, then across the 4 waves through LDS;
//   2. publishes the candidate (key, row, and the row's values) plus, from
//      the owner of row k, row k's values, through device-coherent stores;
//   3. meets the other workgroups at a grid barrier (one relaxed atomic add
//      and a bounded spin; a timeout sets *err and the kernel still ends);
//   4. every workgroup picks the same winner, performs the interchange
//      (the owner of row k takes the pivot row, the owner of the pivot row
//      takes the old row k: SRC/dgetrf2.f:196-200), scales the multipliers by
//      the reciprocal of the pivot or divides when |pivot| < sfmin
//      (SRC/dgetrf2.f:204-210), and applies the rank-1 update to its rows
//      (the A22 update of SRC/dgetrf2.f:245-246 one column at a time);
//   5. rotates its register row left by one so column k+1 sits in slot 0
//      and the finished column k parks in the top slot; after the last
//      column each thread stores its rows as contiguous runs.
// IPIV receives absolute 1-based row numbers; INFO the first exact zero
// pivot (SRC/dgetrf2.f:192,212-214), found first in stream order.
#include "la_common.h"

namespace la {
namespace {

constexpr int NT = 256;
constexpr int GMAX = 128;   // workgroups per panel launch
constexpr int WMAX = 64;    // widest leaf

template <class T> struct Sfmin;
template <> struct Sfmin<double> { static constexpr double v = 0x1p-1022; };
template <> struct Sfmin<float> { static constexpr float v = 0x1p-126f; };

// Exchange area layout (doubles first, then T rows), two parity slots.
struct Xchg {
  double* keys;  // [2][GMAX][2]  (key, row)
  char* rows;    // [2][GMAX][WMAX] of T
  char* rowk;    // [2][WMAX] of T
};
__host__ __device__ inline size_t xchg_bytes_total() {
  return sizeof(double) * 2 * GMAX * 2 + sizeof(double) * 2 * GMAX * WMAX + sizeof(double) * 2 * WMAX;
}

template <class T>
__device__ __forceinline__ void better(double& ka, int& ia, double kb, int ib) {
  if (kb > ka || (kb == ka && ib < ia)) { ka = kb; ia = ib; }
}

template <class T, int W, int R>
__global__ __launch_bounds__(NT) void getf2_panel_kernel(T* __restrict__ Tm, long ldt, int j0, int w,
                                                        int m, int* __restrict__ ipiv,
                                                        int* __restrict__ info,
                                                        unsigned* __restrict__ counter,
                                                        double* __restrict__ xbuf,
                                                        int* __restrict__ err) {
  __shared__ double s_key[NT / 64];
  __shared__ int s_idx[NT / 64];
  __shared__ T s_prow[W];
  __shared__ T s_rowk[W];
  __shared__ int s_piv;
  __shared__ int s_win;

  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int h = m - j0;                   // panel height
  const int kmax = min(h, w);
  const T sfmin = Sfmin<T>::v;

  double* keys = xbuf;
  T* rows = reinterpret_cast<T*>(xbuf + 2 * GMAX * 2);
  T* rowk = reinterpret_cast<T*>(xbuf + 2 * GMAX * 2 + 2 * GMAX * WMAX);

  // Load: thread owns panel rows lr = b*NT*R + q*NT + tid; its w values are
  // a contiguous run of T (T(j0.., j0+lr) = A(j0+lr, j0..)).
  // Register layout: slot c holds panel column (c + k) mod W during step k
  // (every row rotates left once per step), so the pivot column is always
  // slot 0, finished columns collect at the top, and after W rotations in
  // total slot c holds column c again.
  T v[R][W];
  int lr[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    lr[q] = b * NT * R + q * NT + tid;
    const bool ok = lr[q] < h;
    const T* col = Tm + (long)(j0 + lr[q]) * ldt + j0;
#pragma unroll
    for (int c = 0; c < W; ++c) v[q][c] = (ok && c < w) ? col[c] : T(0);
  }

  unsigned target = 0;
  for (int k = 0; k < kmax; ++k) {
    const int par = k & 1;
    // 1. local candidate under the IDAMAX rule.
    double bk = -2.0;
    int bi = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (lr[q] >= k && lr[q] < h) {
        const T x = v[q][0];
        double key;
        if (lr[q] == k) key = (x != x) ? __builtin_huge_val() : (double)fabs(x);
        else key = (x != x) ? -1.0 : (double)fabs(x);
        better<T>(bk, bi, key, lr[q]);
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ok = __shfl_xor(bk, off);
      const int oi = __shfl_xor(bi, off);
      better<T>(bk, bi, ok, oi);
    }
    if (lane == 0) { s_key[wave] = bk; s_idx[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      double kk = s_key[0];
      int ii = s_idx[0];
#pragma unroll
      for (int wv = 1; wv < NT / 64; ++wv) better<T>(kk, ii, s_key[wv], s_idx[wv]);
      s_win = ii;
      st_dev(&keys[(par * GMAX + b) * 2 + 0], kk);
      st_dev(&keys[(par * GMAX + b) * 2 + 1], (double)ii);
    }
    __syncthreads();
    // 2. publish the candidate row and (owner only) row k.
    {
      const int win = s_win;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (lr[q] == win) {
#pragma unroll
          for (int c = 0; c < W; ++c) st_dev(&rows[((size_t)par * GMAX + b) * WMAX + c], v[q][c]);
        }
        if (lr[q] == k) {
#pragma unroll
          for (int c = 0; c < W; ++c) st_dev(&rowk[(size_t)par * WMAX + c], v[q][c]);
        }
      }
    }
    wait_vm();
    __syncthreads();
    // 3. grid barrier.
    target += (unsigned)G;
    if (tid == 0) {
      __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // Bounded wait: ~2^20 polls (about a second); after a timeout anywhere
      // every later wait gives up at once, so the kernel always drains.
      unsigned spins = 0;
      while (ld_dev(counter) < target) {
        __builtin_amdgcn_s_sleep(1);
        if ((++spins & 255u) == 0 && ld_dev(err) != 0) break;
        if (spins > (1u << 20)) { st_dev(err, 1); break; }
      }
    }
    __syncthreads();
    // 4. global winner, pivot row, old row k.
    if (wave == 0) {
      double kk = -3.0;
      int ii = 0x7fffffff, gg = 0;
      for (int g = lane; g < G; g += 64) {
        const double kg = ld_dev(&keys[(par * GMAX + g) * 2 + 0]);
        const int ig = (int)ld_dev(&keys[(par * GMAX + g) * 2 + 1]);
        if (kg > kk || (kg == kk && ig < ii)) { kk = kg; ii = ig; gg = g; }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ok = __shfl_xor(kk, off);
        const int oi = __shfl_xor(ii, off);
        const int og = __shfl_xor(gg, off);
        if (ok > kk || (ok == kk && oi < ii)) { kk = ok; ii = oi; gg = og; }
      }
      for (int c = lane; c < W; c += 64) {
        s_prow[c] = ld_dev(&rows[((size_t)par * GMAX + gg) * WMAX + c]);
        s_rowk[c] = ld_dev(&rowk[(size_t)par * WMAX + c]);
      }
      if (lane == 0) s_piv = ii;
    }
    __syncthreads();
    const int p = s_piv;
    const T piv = s_prow[0];
    if (b == 0 && tid == 0) {
      ipiv[j0 + k] = j0 + p + 1;
      if (piv == T(0)) atomicCAS(info, 0, j0 + k + 1);
    }
    const bool use_recip = fabs(piv) >= sfmin;
    const T rp = T(1) / piv;
    const int lim = W - k;  // slots 1..lim-1 hold the unfinished columns
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (lr[q] >= h) continue;
      if (lr[q] == k) {
#pragma unroll
        for (int c = 0; c < W; ++c) v[q][c] = s_prow[c];
      } else if (lr[q] == p) {
#pragma unroll
        for (int c = 0; c < W; ++c) v[q][c] = s_rowk[c];
      }
      if (lr[q] > k) {
        T l = v[q][0];
        if (piv != T(0)) l = use_recip ? rp * l : l / piv;
        v[q][0] = l;
#pragma unroll
        for (int c = 1; c < W; ++c) {
          const T u = v[q][c] - l * s_prow[c];
          v[q][c] = (c < lim) ? u : v[q][c];
        }
      }
    }
    // 5. rotate: the finished column k moves to the top slot.
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const T t0 = v[q][0];
#pragma unroll
      for (int c = 0; c < W - 1; ++c) v[q][c] = v[q][c + 1];
      v[q][W - 1] = t0;
    }
    // s_prow/s_rowk are rewritten next step only after that step's first
    // two __syncthreads, so no extra barrier is needed here.
  }
  // Complete W rotations so slot c holds column c, then store each row as
  // one contiguous run.
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

int panel_width_for(int h) { return h <= GMAX * NT ? WMAX : WMAX / 2; }

template <class T>
void launch_panel_t(T* Tm, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned* counter,
                    double* xchg, int* err, hipStream_t s) {
  const int h = m - j0;
  if (h <= 0 || w <= 0) return;
  if (w > panel_width_for(h)) return;  // caller contract, checked by the driver
  if (h <= GMAX * NT) {
    const int G = (h + NT - 1) / NT;
    getf2_panel_kernel<T, WMAX, 1><<<G, NT, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, counter, xchg, err);
  } else {
    const int G = (h + 2 * NT - 1) / (2 * NT);
    getf2_panel_kernel<T, WMAX / 2, 2><<<G, NT, 0, s>>>(Tm, ldt, j0, w, m, ipiv, info, counter, xchg, err);
  }
}

}  // namespace

int panel_max_width(int h) { return panel_width_for(h); }
size_t panel_xchg_bytes(int) { return xchg_bytes_total(); }

void launch_getf2_panel_d(double* T, long ldt, int j0, int w, int m, int* ipiv, int* info,
                          unsigned* counter, double* xchg, int* err, hipStream_t s) {
  launch_panel_t<double>(T, ldt, j0, w, m, ipiv, info, counter, xchg, err, s);
}
void launch_getf2_panel_s(float* T, long ldt, int j0, int w, int m, int* ipiv, int* info,
                          unsigned* counter, double* xchg, int* err, hipStream_t s) {
  launch_panel_t<float>(T, ldt, j0, w, m, ipiv, info, counter, xchg, err, s);
}

}  // namespace la
