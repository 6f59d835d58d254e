// aux.hip -- data-movement kernels of the LU path: transposes into and out of
// the factorization workspace, row interchanges (DLASWP, SRC/dlaswp.f:112-189),
// interchange-sequence -> permutation conversion for DGETRS, triangular
// block inverses used by the blocked triangular solves, copies and fills.
#include "la_common.h"

namespace la {
namespace {

// ---- transpose ------------------------------------------------------------
// out(j,i) = cast(in(i,j)), 32x32 tiles through LDS (+1 padding against bank
// conflicts), reads and writes both coalesced.
// skip (optional): nothing is written when *skip != 0 (the factorization's
// write-back after a bounded wait gave up leaves the caller's matrix as it was).
template <class Ti, class To, bool CHECK>
__global__ __launch_bounds__(256) void transpose_kernel(int rows, int cols, const Ti* __restrict__ in,
                                                        long ld_in, To* __restrict__ out, long ld_out,
                                                        int* overflow, const int* __restrict__ skip) {
  __shared__ Ti tile[32][33];
  if (skip && *skip) return;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
  for (int r = 0; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + ty + r;
    if (i < rows && j < cols) tile[ty + r][tx] = in[(long)i + (long)j * ld_in];
  }
  __syncthreads();
  bool ovf = false;
#pragma unroll
  for (int r = 0; r < 32; r += 8) {
    const int j = j0 + tx, i = i0 + ty + r;
    if (i < rows && j < cols) {
      const Ti v = tile[tx][ty + r];
      if (CHECK) {
        // DLAG2S rejects |a| > SLAMCH('O') (SRC/dlag2s.f:135-139).
        if (v < -(Ti)3.4028234663852886e38 || v > (Ti)3.4028234663852886e38) ovf = true;
      }
      out[(long)j + (long)i * ld_out] = (To)v;
    }
  }
  if (CHECK && ovf) atomicOr(overflow, 1);
}

template <class Ti, class To, bool CHECK>
void launch_transpose_t(int rows, int cols, const Ti* in, long ld_in, To* out, long ld_out, int* ovf,
                        hipStream_t s, const int* skip = nullptr) {
  if (rows <= 0 || cols <= 0) return;
  dim3 grid((rows + 31) / 32, (cols + 31) / 32);
  transpose_kernel<Ti, To, CHECK><<<grid, 256, 0, s>>>(rows, cols, in, ld_in, out, ld_out, ovf, skip);
}

// ---- row interchanges in the transposed workspace ---------------------------
// T = A^T: interchanging rows r and p of A is interchanging columns r and p of
// T.  The panel's interchange sequence ipiv[k1..k2) (SRC/dlaswp.f:156-166 order)
// is first composed into (dst, src) pairs over the positions it touches --
// one workgroup walks the sequence once over a position map in LDS -- and the
// columns are then moved in a single pass: each workgroup owns 16 T rows (one
// 128-byte line per T column), stages every source value in LDS, barriers,
// and writes every destination.  Two row ranges (left and right of the panel)
// are served by one launch.
constexpr int SWP_ROWS = 16;      // T rows per workgroup
constexpr int SWP_MAXPAIRS = 512; // 2 * the largest panel width

// Composition of one block of K <= 256 interchanges in parallel.  Steps are
// positions q = b0 .. b0+K-1, step q swapping q with its target p_q >= q
// (the form DGETRF's IPIV always has).  Write c_q for the content of
// position q just before step q.  Before step q position q was last written
// by the latest earlier step u with p_u = q, which left c_u there, so
//   c_q = c_par(q), par(q) = max{u < q : p_u = q}, or q itself if none --
// a forest whose roots are their own contents, collapsed by pointer jumping.
// Position q is final after step q and holds what p_q held before it: c_q
// when p_q = q, else c_prv(q), prv(q) = max{u < q : p_u = p_q}, or p_q itself
// if none.  A target past the block holds c_u of the last step u aiming at
// it.  Every touched position gets exactly one (dst, src) pair.  Called by
// all 256 threads; s_p: the K targets (absolute, 0-based); s_r: 256 ints of
// scratch; pairs go to s_out through the counter *s_cnt.
__device__ void compose_block(int b0, int K, const int* s_p, int* s_r, int* s_out, int* s_cnt) {
  const int t = threadIdx.x;
  const int q = b0 + t;
  int p = q, par = -1, prv = -1;
  bool last = true;
  if (t < K) {
    p = s_p[t];
    // s_p holds -1 past K up to a multiple of 4: never a position
    const int Kr = (K + 3) & ~3;
#pragma unroll 4
    for (int u0 = 0; u0 < Kr; u0 += 4) {
      const int4 pv = *reinterpret_cast<const int4*>(&s_p[u0]);
      const int pu[4] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
      for (int z = 0; z < 4; ++z) {
        const int u = u0 + z;
        if (u < t) {
          if (pu[z] == q) par = u;
          if (pu[z] == p) prv = u;
        } else if (u > t && pu[z] == p) {
          last = false;
        }
      }
    }
    s_r[t] = par >= 0 ? par : t;
  }
  __syncthreads();
  // pointer jumping until every chain has reached its root (at most 8
  // rounds for 256 steps; usually one)
  for (int it = 0; it < 8; ++it) {
    const int cur = t < K ? s_r[t] : 0;
    const int nr = t < K ? s_r[cur] : 0;
    const int moved = __syncthreads_or(nr != cur ? 1 : 0);
    if (!moved) break;
    if (t < K) s_r[t] = nr;
    __syncthreads();
  }
  if (t < K) {
    const int src = p == q ? b0 + s_r[t] : (prv >= 0 ? b0 + s_r[prv] : p);
    if (src != q) {
      const int o = atomicAdd(s_cnt, 1);
      s_out[2 * o] = q;
      s_out[2 * o + 1] = src;
    }
    if (last && p >= b0 + K) {
      const int o = atomicAdd(s_cnt, 1);
      s_out[2 * o] = p;
      s_out[2 * o + 1] = b0 + s_r[t];
    }
  }
  __syncthreads();
}

// pairs[2q] = dst, pairs[2q+1] = src (0-based positions); *npairs = count.
// Positions live in [k1, npos).  Up to 256 interchanges that all satisfy
// p >= q are composed in parallel (compose_block); otherwise one thread walks
// the sequence over a position map (LDS, or global when it does not fit).
template <bool LDSMAP>
__global__ __launch_bounds__(256) void swaps_to_pairs_kernel(int k1, int k2, int npos, const int* __restrict__ ipiv,
                                                             int* __restrict__ pairs, int* __restrict__ npairs,
                                                             int* __restrict__ gmap) {
  extern __shared__ int smem[];
  __shared__ int s_cnt;
  __shared__ __align__(16) int s_ip[SWP_MAXPAIRS / 2];
  __shared__ int s_r[256];
  __shared__ int s_out[2 * SWP_MAXPAIRS];
  const int K = k2 - k1;
  if ((int)threadIdx.x >= K && threadIdx.x < SWP_MAXPAIRS / 2) s_ip[threadIdx.x] = -1;
  // the sequence itself is read once, in parallel; entries outside [k1, npos)
  // are ignored, as a self-interchange
  bool bad = false;
  for (int x = threadIdx.x; x < K && x < SWP_MAXPAIRS / 2; x += 256) {
    int p = ipiv[k1 + x] - 1;
    if (p < k1 || p >= npos) p = k1 + x;
    if (p < k1 + x) bad = true;
    s_ip[x] = p;
  }
  if (threadIdx.x == 0) s_cnt = 0;
  if (__syncthreads_or(bad ? 1 : 0) == 0 && K <= 256) {
    compose_block(k1, K, s_ip, s_r, s_out, &s_cnt);
    const int np = min(s_cnt, SWP_MAXPAIRS);
    for (int x = threadIdx.x; x < 2 * np; x += 256) pairs[x] = s_out[x];
    if (threadIdx.x == 0) *npairs = np;
    return;
  }
  int* cur = LDSMAP ? smem : gmap;
  const int span = npos - k1;
  for (int x = threadIdx.x; x < span; x += 256) cur[x] = k1 + x;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = k1; i < k2; ++i) {
      const int p = i - k1 < SWP_MAXPAIRS / 2 ? s_ip[i - k1] : ipiv[i] - 1;
      if (p != i && p >= k1 && p < npos) {
        const int t = cur[i - k1];
        cur[i - k1] = cur[p - k1];
        cur[p - k1] = t;
      }
    }
  }
  __syncthreads();
  for (int x = threadIdx.x; x < span; x += 256) {
    const int src = cur[x];
    if (src != k1 + x) {
      const int q = atomicAdd(&s_cnt, 1);
      if (q < SWP_MAXPAIRS) {
        pairs[2 * q] = k1 + x;
        pairs[2 * q + 1] = src;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *npairs = min(s_cnt, SWP_MAXPAIRS);
}

template <class T>
__global__ __launch_bounds__(256) void apply_pairs_cols_kernel(T* __restrict__ Tm, long ldt, int r0, int r1, int r2,
                                                               int r3, const int* __restrict__ pairs,
                                                               const int* __restrict__ npairs, int nblk,
                                                               int nskip) {
  // nskip: blocks dispatched to XCDs 0 .. nskip-1 (blockIdx % 8) leave at once
  int blk = blockIdx.x;
  if (nskip) {
    if (blk % 8 < nskip) return;
    blk = (blk / 8) * (8 - nskip) + (blk % 8 - nskip);
  }
  if (blk >= nblk) return;
  const int np = *npairs;
  if (np <= 0) return;
  // this block's 16 rows: first from [r0, r1), then from [r2, r3)
  const int n1 = (r1 - r0 + SWP_ROWS - 1) / SWP_ROWS;
  const long c0 = blk < n1 ? (long)r0 + (long)blk * SWP_ROWS : (long)r2 + (long)(blk - n1) * SWP_ROWS;
  const long cend = blk < n1 ? r1 : r3;
  const int cl = threadIdx.x & (SWP_ROWS - 1);
  const int qg = threadIdx.x / SWP_ROWS;  // 16 pair lanes
  const long c = c0 + cl;
  const bool ok = c < cend;
  // every pair of the thread (at most SWP_MAXPAIRS / 16), then every source
  // in registers, all in flight together; the loads complete and one barrier
  // (a destination may be another pair's source); then every destination
  constexpr int QS = 256 / SWP_ROWS, U = SWP_MAXPAIRS / QS;
  int src[U], dst[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int q = min(qg + QS * u, np - 1);
    src[u] = pairs[2 * q + 1];
    dst[u] = pairs[2 * q];
  }
  wait_vm();  // all indices here: the source loads below then go out back to back
  T v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) v[u] = (ok && qg + QS * u < np) ? Tm[c + (long)src[u] * ldt] : T(0);
  wait_vm();
  __syncthreads();
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (ok && qg + QS * u < np) Tm[c + (long)dst[u] * ldt] = v[u];
}

// Everything between two leaves of a panel in one single-workgroup launch
// (a single workgroup never waits for a share of CUs held by a concurrent
// trailing update, which a many-workgroup launch does):
//   1. the leaf's interchanges ipiv[jj..jj+w) (targets in [jj, m), each at
//      or after its step, as the panel kernels produce them) composed into
//      pairs (compose_block);
//   2. the pairs applied to the panel's other T rows, [j, jj) and
//      [jj+w, jend) (SRC/dgetrf2.f:234-236 on the columns left and right of
//      the leaf), 64 T rows at a time through registers;
//   3. the in-panel DTRSM of SRC/dgetrf2.f:240-241 on the columns right of
//      the leaf: each is solved against the leaf's unit lower triangle by one
//      thread in the reference's order (BLAS/SRC/dtrsm.f:177-190: for k, skip
//      a zero x_k, else x_i -= x_k L(i,k) for i > k).
template <class T>
__global__ __launch_bounds__(256) void panel_boundary_kernel(T* __restrict__ Tm, long ldt, int j, int jj, int w,
                                                             int jend, int m, const int* __restrict__ ipiv) {
  constexpr int W = 64;
  __shared__ __align__(16) int s_p[256];
  __shared__ int s_r[256];
  __shared__ int s_out[2 * SWP_MAXPAIRS];
  __shared__ int s_cnt;
  __shared__ T sL[W][W + 1];
  const int tid = threadIdx.x;
  s_p[tid] = -1;
  if (tid < w) {
    int p = ipiv[jj + tid] - 1;
    if (p < jj + tid || p >= m) p = jj + tid;  // not of the panel's form: left alone
    s_p[tid] = p;
  }
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  compose_block(jj, w, s_p, s_r, s_out, &s_cnt);
  const int np = min(s_cnt, 2 * W);
  const int n1 = jj - j, rest = jend - jj - w;
  const int tot = n1 + rest;
  // The leaf's triangle (rows [jj, jj+w) are not interchanged here) is read
  // together with the first chunk of the interchange: L(i,k) = A(jj+i, jj+k)
  // = T(jj+k, jj+i) -> sL[k][i], zero outside the leaf.
  T lv[W * W / 256];
  if (rest > 0) {
#pragma unroll
    for (int u = 0; u < W * W / 256; ++u) {
      const int e = tid + 256 * u;
      const int k = e % W, i = e / W;
      lv[u] = (k < w && i < w) ? Tm[(long)(jj + k) + (long)(jj + i) * ldt] : T(0);
    }
  }
  // 64 T rows at a time: every source value of the chunk in registers
  // (32 per thread), one barrier, every destination (chunks touch disjoint
  // rows, so the next chunk's loads need no barrier after these stores).
  // Thread tid moves T row r0 + tid % 64 of pairs tid / 64 + 4u, u < 32:
  // the same pairs in every chunk, so their indices are read once.
  constexpr int CH = 64;
  constexpr int PER = 2 * W * CH / 256;
  static_assert(256 % CH == 0 && PER * (256 / CH) == 2 * W, "pair lanes");
  int psrc[PER], pdst[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int q = tid / CH + (256 / CH) * u;
    psrc[u] = q < np ? s_out[2 * q + 1] : 0;
    pdst[u] = q < np ? s_out[2 * q] : 0;
  }
  for (int r0 = 0; r0 < tot; r0 += CH) {
    const int ri = r0 + tid % CH;
    const bool okr = ri < tot;
    const long trow = ri < n1 ? (long)j + ri : (long)jj + w + (ri - n1);
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = tid / CH + (256 / CH) * u;
      v[u] = (okr && q < np) ? Tm[trow + (long)psrc[u] * ldt] : T(0);
    }
    wait_vm();  // every source read before any destination is written
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = tid / CH + (256 / CH) * u;
      if (okr && q < np) Tm[trow + (long)pdst[u] * ldt] = v[u];
    }
  }
  if (rest <= 0) return;
#pragma unroll
  for (int u = 0; u < W * W / 256; ++u) {
    const int e = tid + 256 * u;
    sL[e % W][e / W] = lv[u];
  }
  __syncthreads();
  if (tid < rest) {
    const long trow = (long)jj + w + tid;
    T x[W];
#pragma unroll
    for (int i = 0; i < W; ++i) x[i] = i < w ? Tm[trow + (long)(jj + i) * ldt] : T(0);
    // row k+1 of the triangle is read into registers before row k is
    // applied, so its LDS reads are in flight while row k's FMAs run (read
    // inside the branch, only two reads were in flight at a time)
    T lc[W], ln[W];
#pragma unroll
    for (int i = 1; i < W; ++i) lc[i] = sL[0][i];
#pragma unroll
    for (int k = 0; k < W; ++k) {
#pragma unroll
      for (int i = k + 2; i < W; ++i) ln[i] = sL[k + 1][i];
      if (x[k] != T(0)) {
#pragma unroll
        for (int i = k + 1; i < W; ++i) x[i] = x[i] - x[k] * lc[i];
      }
#pragma unroll
      for (int i = k + 2; i < W; ++i) lc[i] = ln[i];
    }
#pragma unroll
    for (int i = 0; i < W; ++i)
      if (i < w) Tm[trow + (long)(jj + i) * ldt] = x[i];
  }
}

// Row interchanges of a column-major matrix through the same pairs:
// B(dst, j) <- B_old(src, j); one workgroup per column block of 16.
template <class T>
__global__ __launch_bounds__(256) void apply_pairs_rows_kernel(T* __restrict__ B, long ldb, int ncols,
                                                               const int* __restrict__ pairs,
                                                               const int* __restrict__ npairs) {
  const int np = *npairs;
  if (np <= 0) return;
  const int j = blockIdx.x * SWP_ROWS + (threadIdx.x & (SWP_ROWS - 1));
  const int qg = threadIdx.x / SWP_ROWS;
  const bool ok = j < ncols;
  // as apply_pairs_cols_kernel: pairs, then sources in registers, one
  // barrier, every destination
  constexpr int QS = 256 / SWP_ROWS, U = SWP_MAXPAIRS / QS;
  int src[U], dst[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int q = min(qg + QS * u, np - 1);
    src[u] = pairs[2 * q + 1];
    dst[u] = pairs[2 * q];
  }
  wait_vm();
  T v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) v[u] = (ok && qg + QS * u < np) ? B[(long)src[u] + (long)j * ldb] : T(0);
  wait_vm();
  __syncthreads();
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (ok && qg + QS * u < np) B[(long)dst[u] + (long)j * ldb] = v[u];
}

// ---- DLASWP on a column-major matrix (row interchanges) ----------------------
// Used on the right-hand sides of DGETRS (SRC/dgetrs.f:187,217) and by the
// dlaswp_ entry point.  The interchange sequence is first turned into a
// permutation (perm[r] = the row whose content ends at row r), walked once
// in the reference's order (SRC/dlaswp.f:138-150: K1..K2 for INCX > 0, K2..K1
// for INCX < 0) over a position map held in LDS; the rows are then gathered
// through it in parallel.
constexpr int PERM_LDS_MAX = 32768;
constexpr int PERM_CHUNK = 1024;

template <bool LDSMAP>
__global__ __launch_bounds__(256) void swaps_to_perm_kernel(int nrows, int k1, int k2,
                                                            const int* __restrict__ ipiv, int incx,
                                                            int* __restrict__ perm) {
  extern __shared__ int smem[];
  int* sp = smem;                                  // PERM_CHUNK staged pivots
  int* cur = LDSMAP ? smem + PERM_CHUNK : perm;    // position map
  for (int r = threadIdx.x; r < nrows; r += 256) cur[r] = r;
  __syncthreads();
  const int nk = k2 - k1;
  for (int c0 = 0; c0 < nk; c0 += PERM_CHUNK) {
    const int cn = min(PERM_CHUNK, nk - c0);
    // chunk c0 in processing order: forward k1+c0.., reverse k2-1-c0 downwards
    for (int t = threadIdx.x; t < cn; t += 256) {
      const int i = incx > 0 ? k1 + c0 + t : k2 - 1 - c0 - t;
      sp[t] = ipiv[i] - 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int t = 0; t < cn; ++t) {
        const int i = incx > 0 ? k1 + c0 + t : k2 - 1 - c0 - t;
        const int p = sp[t];
        if (p != i && p >= 0 && p < nrows) {  // out-of-range entries are ignored
          const int a = cur[i];
          cur[i] = cur[p];
          cur[p] = a;
        }
      }
    }
    __syncthreads();
  }
  if (LDSMAP)
    for (int r = threadIdx.x; r < nrows; r += 256) perm[r] = cur[r];
}

// The same permutation for the sequences DGETRF produces (every target at or
// after its step), in blocks of 256 interchanges: each block is composed in
// parallel (compose_block) and applied to the position map as one gather;
// a block breaking that form is walked serially.  INCX < 0 applies the
// interchanges in reverse order, i.e. the inverse permutation of the
// forward sequence: the forward map is built and inverted.
__global__ __launch_bounds__(256) void swaps_to_perm_blocked_kernel(int nrows, int k1, int k2,
                                                                    const int* __restrict__ ipiv, int incx,
                                                                    int* __restrict__ perm) {
  extern __shared__ int cur[];  // nrows ints
  __shared__ __align__(16) int s_p[256];
  __shared__ int s_r[256];
  __shared__ int s_out[2 * SWP_MAXPAIRS];
  __shared__ int s_cnt;
  for (int r = threadIdx.x; r < nrows; r += 256) cur[r] = r;
  for (int b0 = k1; b0 < k2; b0 += 256) {
    const int K = min(256, k2 - b0);
    bool bad = false;
    s_p[threadIdx.x] = -1;
    if ((int)threadIdx.x < K) {
      const int q = b0 + threadIdx.x;
      int p = ipiv[q] - 1;
      if (p < 0 || p >= nrows) p = q;  // ignored, as a self-interchange
      if (p < q) bad = true;
      s_p[threadIdx.x] = p;
    }
    if (threadIdx.x == 0) s_cnt = 0;
    if (__syncthreads_or(bad ? 1 : 0) == 0) {
      compose_block(b0, K, s_p, s_r, s_out, &s_cnt);
      const int np = s_cnt;
      int d[2], v[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int o = threadIdx.x + 256 * h;
        d[h] = o < np ? s_out[2 * o] : -1;
        v[h] = o < np ? cur[s_out[2 * o + 1]] : 0;
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (d[h] >= 0) cur[d[h]] = v[h];
    } else if (threadIdx.x == 0) {
      for (int t = 0; t < K; ++t) {
        const int i = b0 + t, p = s_p[t];
        if (p != i) {
          const int a = cur[i];
          cur[i] = cur[p];
          cur[p] = a;
        }
      }
    }
    __syncthreads();
  }
  if (incx > 0) {
    for (int r = threadIdx.x; r < nrows; r += 256) perm[r] = cur[r];
  } else {
    for (int r = threadIdx.x; r < nrows; r += 256) perm[cur[r]] = r;
  }
}

// The same permutation in two launches, for long sequences: every block of
// 256 interchanges is composed at once, one workgroup per block
// (compose_blocks_kernel: block b's pairs at gp + b*PERM_PB, its pair count
// -- or -1 for a block not of DGETRF's form -- at gp[b*PERM_PB +
// 2*SWP_MAXPAIRS]); then one workgroup applies the blocks to the position
// map in order (swaps_to_perm_pairs_kernel), reading block b+1's pairs
// while it applies block b's, so a block costs an LDS gather and two
// barriers instead of a composition.
constexpr int PERM_PB = 2 * SWP_MAXPAIRS + 1;

__global__ __launch_bounds__(256) void compose_blocks_kernel(int nrows, int k1, int k2,
                                                             const int* __restrict__ ipiv, int* __restrict__ gp) {
  __shared__ __align__(16) int s_p[256];
  __shared__ int s_r[256];
  __shared__ int s_out[2 * SWP_MAXPAIRS];
  __shared__ int s_cnt;
  const int b0 = k1 + 256 * (int)blockIdx.x;
  if (b0 >= k2) return;
  const int K = min(256, k2 - b0);
  int* out = gp + (long)blockIdx.x * PERM_PB;
  bool bad = false;
  s_p[threadIdx.x] = -1;
  if ((int)threadIdx.x < K) {
    const int q = b0 + threadIdx.x;
    int p = ipiv[q] - 1;
    if (p < 0 || p >= nrows) p = q;  // ignored, as a self-interchange
    if (p < q) bad = true;
    s_p[threadIdx.x] = p;
  }
  if (threadIdx.x == 0) s_cnt = 0;
  if (__syncthreads_or(bad ? 1 : 0) != 0) {
    if (threadIdx.x == 0) out[2 * SWP_MAXPAIRS] = -1;
    return;
  }
  compose_block(b0, K, s_p, s_r, s_out, &s_cnt);
  const int np = s_cnt;
  for (int o = threadIdx.x; o < 2 * np; o += 256) out[o] = s_out[o];
  if (threadIdx.x == 0) out[2 * SWP_MAXPAIRS] = np;
}

__global__ __launch_bounds__(256) void swaps_to_perm_pairs_kernel(int nrows, int k1, int k2,
                                                                  const int* __restrict__ ipiv, int incx,
                                                                  const int* __restrict__ gp,
                                                                  int* __restrict__ perm) {
  extern __shared__ int cur[];  // nrows ints
  const int tid = threadIdx.x;
  for (int r = tid; r < nrows; r += 256) cur[r] = r;
  const int nb = (k2 - k1 + 255) / 256;
  // block b's pair slots o = tid and tid + 256 (at most 2 * 256 pairs)
  int cnt = gp[2 * SWP_MAXPAIRS], d[2], sr[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    d[h] = gp[2 * (tid + 256 * h)];
    sr[h] = gp[2 * (tid + 256 * h) + 1];
  }
  __syncthreads();
  for (int b = 0; b < nb; ++b) {
    // the next block's pairs, in flight while this block is applied
    const int bn = min(b + 1, nb - 1);
    const int* gn = gp + (long)bn * PERM_PB;
    const int cntn = gn[2 * SWP_MAXPAIRS];
    int dn[2], srn[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      dn[h] = gn[2 * (tid + 256 * h)];
      srn[h] = gn[2 * (tid + 256 * h) + 1];
    }
    if (cnt >= 0) {
      int v[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) v[h] = tid + 256 * h < cnt ? cur[sr[h]] : 0;
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (tid + 256 * h < cnt) cur[d[h]] = v[h];
    } else if (tid == 0) {
      const int b0 = k1 + 256 * b, K = min(256, k2 - b0);
      for (int t = 0; t < K; ++t) {
        const int i = b0 + t;
        int p = ipiv[i] - 1;
        if (p < 0 || p >= nrows) p = i;
        if (p != i) {
          const int a = cur[i];
          cur[i] = cur[p];
          cur[p] = a;
        }
      }
    }
    __syncthreads();
    cnt = cntn;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      d[h] = dn[h];
      sr[h] = srn[h];
    }
  }
  if (incx > 0) {
    for (int r = tid; r < nrows; r += 256) perm[r] = cur[r];
  } else {
    for (int r = tid; r < nrows; r += 256) perm[cur[r]] = r;
  }
}

// dst(r,j) = src(perm[r], j) for r < nrows, j < ncols.
template <class T>
__global__ __launch_bounds__(256) void gather_rows_kernel(int nrows, int ncols, const int* __restrict__ perm,
                                                          const T* __restrict__ src, long lds,
                                                          T* __restrict__ dst, long ldd, const int* __restrict__ skip) {
  if (skip && *skip) return;
  const long tot = (long)nrows * ncols;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const int r = (int)(e % nrows), j = (int)(e / nrows);
    dst[(long)r + (long)j * ldd] = src[(long)perm[r] + (long)j * lds];
  }
}

// The same rows moved in place, one column at a time through LDS:
// A(r,j) <- A_old(perm[r], j) for r < nrows.  The rows the permutation moves
// are read into LDS and then written from their sources, so each column is
// read and written once (a gather into scratch and a copy back read and write
// it twice).  A fixed point r (perm[r] == r) is neither read nor written: a
// bijection takes every source of a moved row from the moved rows.  Thread
// tid holds perm of rows tid + 256u, u < U, in registers over all the columns
// of its workgroup, and loads the next column while it writes this one.
constexpr int PERMUTE_LDS_BYTES = 131072;

template <class T, int U>
__global__ __launch_bounds__(256) void permute_rows_lds_kernel(int nrows, int ncols, const int* __restrict__ perm,
                                                               T* __restrict__ A, long lda) {
  __shared__ __align__(16) T col[PERMUTE_LDS_BYTES / sizeof(T)];
  const int tid = threadIdx.x;
  int j = blockIdx.x;
  if (j >= ncols) return;
  int pr[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int r = tid + 256 * u;
    pr[u] = r < nrows ? min(max(perm[r], 0), nrows - 1) : r;
  }
  T v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int r = tid + 256 * u;
    v[u] = pr[u] != r ? A[(long)r + (long)j * lda] : T(0);
  }
  for (; j < ncols; j += gridDim.x) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = tid + 256 * u;
      if (pr[u] != r) col[r] = v[u];
    }
    __syncthreads();
    const int jn = j + gridDim.x;
    if (jn < ncols) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = tid + 256 * u;
        v[u] = pr[u] != r ? A[(long)r + (long)jn * lda] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = tid + 256 * u;
      if (pr[u] != r) A[(long)r + (long)j * lda] = col[pr[u]];
    }
    __syncthreads();  // every read of col done before the next column's values go in
  }
}

// ---- triangular block inverses -------------------------------------------
// W = inv(tri(U)) for nb x nb blocks, nb <= 256; blockIdx.x = column c of W,
// blockIdx.y = which diagonal block (batched along the diagonal with stride
// nb: block d starts at U + d*nb*(ldu+1), W + d*nb*ldw).  One thread per row.
// UPPER: solve tri(U) x = e_c by backward substitution (column-oriented, the
// loop order of BLAS/SRC/dtrsm.f:277-288 on a single right-hand side);
// lower: forward substitution (dtrsm.f:289-300).  UNIT: unit diagonal.
template <class T, bool UPPER, bool UNIT>
__global__ __launch_bounds__(256) void trtri_kernel(int n, int nb, const T* __restrict__ U, long ldu,
                                                    T* __restrict__ W, long ldw) {
  // the solved value of step t goes through xs[t & 1]: one barrier per step
  // (a slot is rewritten two steps later, after every thread has read it)
  __shared__ T xs[2];
  const int c = blockIdx.x;
  const int d = blockIdx.y;
  const int b0 = d * nb;
  const int bs = min(nb, n - b0);
  if (c >= bs) return;
  const T* Ub = U + (long)b0 * (ldu + 1);
  T* Wb = W + (long)d * nb * ldw;  // W block d: nb x nb at column offset d*nb
  const int s = threadIdx.x;
  T x = (s == c) ? T(1) : T(0);
  // the steps in groups of D; each group's triangle entries are loaded while
  // the group before it runs (two register sets in turn), so the loads wait
  // behind D barrier steps instead of in front of them
  constexpr int D = 8;
  T ua[D], ub[D];
  // entries of the group starting at step t0 (descending for UPPER), read
  // from clamped in-bounds addresses without a branch (so that the first use
  // waits only for its own load); the update below uses an entry only for
  // the rows it belongs to
  const int sc = min(s, bs - 1);
  auto load = [&](T (&u)[D], int t0) {
#pragma unroll
    for (int e = 0; e < D; ++e) {
      const int t = min(max(UPPER ? t0 - e : t0 + e, 0), bs - 1);
      u[e] = Ub[(long)sc + (long)t * ldu];
    }
  };
  auto run = [&](const T (&u)[D], int t0) {
#pragma unroll
    for (int e = 0; e < D; ++e) {
      const int t = UPPER ? t0 - e : t0 + e;
      if (UPPER ? t < 0 : t >= bs) break;
      if (s == t) {
        if (!UNIT) x = x / Ub[(long)t + (long)t * ldu];
        xs[t & 1] = x;
      }
      __syncthreads();
      const T xt = xs[t & 1];
      if (UPPER ? s < t : (s > t && s < bs)) x = x - xt * u[e];
    }
  };
  if (UPPER) {
    load(ua, c);
    for (int t0 = c; t0 >= 0; t0 -= 2 * D) {
      load(ub, t0 - D);
      run(ua, t0);
      if (t0 - D < 0) break;
      load(ua, t0 - 2 * D);
      run(ub, t0 - D);
    }
  } else {
    load(ua, c);
    for (int t0 = c; t0 < bs; t0 += 2 * D) {
      load(ub, t0 + D);
      run(ua, t0);
      if (t0 + D >= bs) break;
      load(ua, t0 + 2 * D);
      run(ub, t0 + D);
    }
  }
  if (s < bs) Wb[(long)s + (long)c * ldw] = x;
}

// ---- copies and fills -----------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void copy_kernel(int rows, int cols, const T* __restrict__ src, long lds,
                                                   T* __restrict__ dst, long ldd, int nskip,
                                                   const int* __restrict__ skip) {
  if (skip && *skip) return;
  // nskip: blocks dispatched to XCDs 0 .. nskip-1 (blockIdx % 8) leave at once
  long lb = blockIdx.x, ng = gridDim.x;
  if (nskip) {
    if ((int)(blockIdx.x % 8) < nskip) return;
    lb = (long)(blockIdx.x / 8) * (8 - nskip) + (blockIdx.x % 8 - nskip);
    ng = (long)(gridDim.x / 8) * (8 - nskip);
  }
  const long tot = (long)rows * cols;
  for (long e = lb * 256 + threadIdx.x; e < tot; e += ng * 256) {
    const long i = e % rows, j = e / rows;
    dst[i + j * ldd] = src[i + j * lds];
  }
}
__global__ void fold_err_kernel(const int* __restrict__ err, int* __restrict__ info, int* __restrict__ sticky) {
  if (threadIdx.x == 0 && *err) {
    if (info) *info = LA_ERR_SYNC;
    if (sticky) atomicOr(sticky, 1);
  }
}

// x[i] uniform in [-1, 1) from a counter-based hash of (seed, i) (a
// splitmix64 finalizer): test and tuning inputs made on the device.
__global__ __launch_bounds__(256) void fill_random_kernel(double* __restrict__ x, size_t n, unsigned long long seed) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    unsigned long long z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z = z ^ (z >> 31);
    x[i] = (double)(z >> 11) * 0x1p-52 - 1.0;
  }
}

template <class T>
__global__ __launch_bounds__(256) void fill_kernel(T* __restrict__ p, size_t n, T v) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) p[e] = v;
}

inline int grid_for(long work) {
  long g = (work + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

void launch_transpose_d(int rows, int cols, const double* in, long ld_in, double* out, long ld_out,
                        hipStream_t s, const int* skip) {
  launch_transpose_t<double, double, false>(rows, cols, in, ld_in, out, ld_out, nullptr, s, skip);
}
void launch_transpose_s(int rows, int cols, const float* in, long ld_in, float* out, long ld_out,
                        hipStream_t s, const int* skip) {
  launch_transpose_t<float, float, false>(rows, cols, in, ld_in, out, ld_out, nullptr, s, skip);
}
void launch_transpose_d2s(int rows, int cols, const double* in, long ld_in, float* out, long ld_out,
                          int* overflow, hipStream_t s) {
  launch_transpose_t<double, float, true>(rows, cols, in, ld_in, out, ld_out, overflow, s);
}

int lds_limit() {
  static int lds_max = -1;
  if (lds_max < 0) {
    int dev = 0, v = 65536;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
    lds_max = v;
  }
  return lds_max;
}

int swp_max_pairs() { return SWP_MAXPAIRS; }

// (dst, src) pairs of the interchange sequence ipiv[k1..k2) over positions
// [k1, npos); gmap: npos ints of scratch used when the map does not fit LDS.
void launch_swaps_to_pairs(int k1, int k2, int npos, const int* ipiv, int* pairs, int* npairs, int* gmap,
                           hipStream_t s) {
  const size_t need = sizeof(int) * (size_t)(npos - k1) + 8192;  // + the kernel's static LDS
  if (need <= (size_t)lds_limit())
    swaps_to_pairs_kernel<true><<<1, 256, sizeof(int) * (size_t)(npos - k1), s>>>(k1, k2, npos, ipiv, pairs, npairs, gmap);
  else
    swaps_to_pairs_kernel<false><<<1, 256, 0, s>>>(k1, k2, npos, ipiv, pairs, npairs, gmap);
}
template <class T>
static void apply_cols_t(T* Tm, long ldt, int r0, int r1, int r2, int r3, const int* pairs, const int* npairs,
                         hipStream_t s, int nskip) {
  if (r1 < r0) r1 = r0;
  if (r3 < r2) r3 = r2;
  const int nb = (r1 - r0 + SWP_ROWS - 1) / SWP_ROWS + (r3 - r2 + SWP_ROWS - 1) / SWP_ROWS;
  const int grid = nskip ? (nb + (8 - nskip) - 1) / (8 - nskip) * 8 : nb;
  if (nb > 0)
    apply_pairs_cols_kernel<T><<<grid, 256, 0, s>>>(Tm, ldt, r0, r1, r2, r3, pairs, npairs, nb, nskip);
}
void launch_apply_pairs_cols_d(double* T, long ldt, int r0, int r1, int r2, int r3, const int* pairs,
                               const int* npairs, hipStream_t s, int nskip) {
  apply_cols_t<double>(T, ldt, r0, r1, r2, r3, pairs, npairs, s, nskip);
}
void launch_apply_pairs_cols_s(float* T, long ldt, int r0, int r1, int r2, int r3, const int* pairs,
                               const int* npairs, hipStream_t s, int nskip) {
  apply_cols_t<float>(T, ldt, r0, r1, r2, r3, pairs, npairs, s, nskip);
}
void launch_panel_boundary_d(double* T, long ldt, int j, int jj, int w, int jend, int m, const int* ipiv,
                             hipStream_t s) {
  panel_boundary_kernel<double><<<1, 256, 0, s>>>(T, ldt, j, jj, w, jend, m, ipiv);
}
void launch_panel_boundary_s(float* T, long ldt, int j, int jj, int w, int jend, int m, const int* ipiv,
                             hipStream_t s) {
  panel_boundary_kernel<float><<<1, 256, 0, s>>>(T, ldt, j, jj, w, jend, m, ipiv);
}
void launch_apply_pairs_rows_d(double* B, long ldb, int ncols, const int* pairs, const int* npairs, hipStream_t s) {
  if (ncols > 0) apply_pairs_rows_kernel<double><<<(ncols + SWP_ROWS - 1) / SWP_ROWS, 256, 0, s>>>(B, ldb, ncols, pairs, npairs);
}
void launch_apply_pairs_rows_s(float* B, long ldb, int ncols, const int* pairs, const int* npairs, hipStream_t s) {
  if (ncols > 0) apply_pairs_rows_kernel<float><<<(ncols + SWP_ROWS - 1) / SWP_ROWS, 256, 0, s>>>(B, ldb, ncols, pairs, npairs);
}

size_t perm_pairs_ints(int k1, int k2) { return k2 > k1 ? (size_t)((k2 - k1 + 255) / 256) * PERM_PB : 0; }

// perm (nrows ints) <- permutation of the interchange sequence; see above.
// gp: perm_pairs_ints(k1, k2) ints of device scratch, or null (then the
// blocks are composed one after another in one workgroup).
void launch_swaps_to_perm(int nrows, int k1, int k2, const int* ipiv, int incx, int* perm, hipStream_t s,
                          int* gp) {
  if (k1 >= 0 && k2 <= nrows && sizeof(int) * (size_t)nrows + 8192 <= (size_t)lds_limit()) {
    if (gp && k2 > k1) {
      const int nb = (k2 - k1 + 255) / 256;
      compose_blocks_kernel<<<nb, 256, 0, s>>>(nrows, k1, k2, ipiv, gp);
      swaps_to_perm_pairs_kernel<<<1, 256, sizeof(int) * (size_t)nrows, s>>>(nrows, k1, k2, ipiv, incx, gp, perm);
      return;
    }
    swaps_to_perm_blocked_kernel<<<1, 256, sizeof(int) * (size_t)nrows, s>>>(nrows, k1, k2, ipiv, incx, perm);
    return;
  }
  const size_t need = sizeof(int) * (PERM_CHUNK + (size_t)nrows);
  if (nrows <= PERM_LDS_MAX && need <= (size_t)lds_limit()) {
    const size_t sh = sizeof(int) * (PERM_CHUNK + (size_t)nrows);
    swaps_to_perm_kernel<true><<<1, 256, sh, s>>>(nrows, k1, k2, ipiv, incx, perm);
  } else {
    swaps_to_perm_kernel<false><<<1, 256, sizeof(int) * PERM_CHUNK, s>>>(nrows, k1, k2, ipiv, incx, perm);
  }
}
void launch_gather_rows_d(int nrows, int ncols, const int* perm, const double* src, long lds, double* dst,
                          long ldd, hipStream_t s, const int* skip) {
  if (nrows > 0 && ncols > 0)
    gather_rows_kernel<double><<<grid_for((long)nrows * ncols), 256, 0, s>>>(nrows, ncols, perm, src, lds, dst, ldd,
                                                                             skip);
}
void launch_gather_rows_s(int nrows, int ncols, const int* perm, const float* src, long lds, float* dst,
                          long ldd, hipStream_t s, const int* skip) {
  if (nrows > 0 && ncols > 0)
    gather_rows_kernel<float><<<grid_for((long)nrows * ncols), 256, 0, s>>>(nrows, ncols, perm, src, lds, dst, ldd,
                                                                            skip);
}

int permute_rows_max(int elem_bytes) { return PERMUTE_LDS_BYTES / elem_bytes; }

template <class T>
static bool launch_permute_rows_t(int nrows, int ncols, const int* perm, T* A, long lda, hipStream_t s) {
  if (nrows <= 0 || ncols <= 0 || nrows > permute_rows_max((int)sizeof(T))) return false;
  const int grid = std::min(ncols, 512);
  if (nrows <= 2048) {
    permute_rows_lds_kernel<T, 8><<<grid, 256, 0, s>>>(nrows, ncols, perm, A, lda);
  } else if (nrows <= 4096) {
    permute_rows_lds_kernel<T, 16><<<grid, 256, 0, s>>>(nrows, ncols, perm, A, lda);
  } else if (nrows <= 8192) {
    permute_rows_lds_kernel<T, 32><<<grid, 256, 0, s>>>(nrows, ncols, perm, A, lda);
  } else if (nrows <= 16384) {
    permute_rows_lds_kernel<T, 64><<<grid, 256, 0, s>>>(nrows, ncols, perm, A, lda);
  } else {
    if constexpr (sizeof(T) == 4) permute_rows_lds_kernel<T, 128><<<grid, 256, 0, s>>>(nrows, ncols, perm, A, lda);
  }
  return true;
}
bool launch_permute_rows_d(int nrows, int ncols, const int* perm, double* A, long lda, hipStream_t s) {
  return launch_permute_rows_t<double>(nrows, ncols, perm, A, lda, s);
}
bool launch_permute_rows_s(int nrows, int ncols, const int* perm, float* A, long lda, hipStream_t s) {
  return launch_permute_rows_t<float>(nrows, ncols, perm, A, lda, s);
}

template <class T>
static void launch_trtri_t(bool upper, bool unit, int n, int nb, const T* U, long ldu, T* W, long ldw,
                           hipStream_t s) {
  if (n <= 0) return;
  const int nblk = (n + nb - 1) / nb;
  dim3 grid(nb, nblk);
  if (upper && unit) trtri_kernel<T, true, true><<<grid, 256, 0, s>>>(n, nb, U, ldu, W, ldw);
  else if (upper) trtri_kernel<T, true, false><<<grid, 256, 0, s>>>(n, nb, U, ldu, W, ldw);
  else if (unit) trtri_kernel<T, false, true><<<grid, 256, 0, s>>>(n, nb, U, ldu, W, ldw);
  else trtri_kernel<T, false, false><<<grid, 256, 0, s>>>(n, nb, U, ldu, W, ldw);
}
void launch_trtri_blocks_d(bool upper, bool unit, int n, int nb, const double* U, long ldu, double* W,
                           long ldw, hipStream_t s) {
  launch_trtri_t<double>(upper, unit, n, nb, U, ldu, W, ldw, s);
}
void launch_trtri_blocks_s(bool upper, bool unit, int n, int nb, const float* U, long ldu, float* W,
                           long ldw, hipStream_t s) {
  launch_trtri_t<float>(upper, unit, n, nb, U, ldu, W, ldw, s);
}
void launch_trtri_uu_d(int nb, const double* U, long ldu, double* W, long ldw, hipStream_t s) {
  launch_trtri_t<double>(true, true, nb, nb, U, ldu, W, ldw, s);
}
void launch_trtri_uu_s(int nb, const float* U, long ldu, float* W, long ldw, hipStream_t s) {
  launch_trtri_t<float>(true, true, nb, nb, U, ldu, W, ldw, s);
}
void launch_trtri_un_d(int nb, const double* U, long ldu, double* W, long ldw, hipStream_t s) {
  launch_trtri_t<double>(true, false, nb, nb, U, ldu, W, ldw, s);
}
void launch_trtri_un_s(int nb, const float* U, long ldu, float* W, long ldw, hipStream_t s) {
  launch_trtri_t<float>(true, false, nb, nb, U, ldu, W, ldw, s);
}

static int copy_grid(long work, int nskip) {
  const int g = grid_for(work);
  return nskip ? (g + (8 - nskip) - 1) / (8 - nskip) * 8 : g;
}
void launch_copy_d(int rows, int cols, const double* src, long lds, double* dst, long ldd, hipStream_t s,
                   int nskip, const int* skip) {
  if (rows > 0 && cols > 0)
    copy_kernel<double><<<copy_grid((long)rows * cols, nskip), 256, 0, s>>>(rows, cols, src, lds, dst, ldd, nskip,
                                                                            skip);
}
void launch_copy_s(int rows, int cols, const float* src, long lds, float* dst, long ldd, hipStream_t s,
                   int nskip, const int* skip) {
  if (rows > 0 && cols > 0)
    copy_kernel<float><<<copy_grid((long)rows * cols, nskip), 256, 0, s>>>(rows, cols, src, lds, dst, ldd, nskip,
                                                                           skip);
}
void launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s) {
  if (n) fill_kernel<unsigned><<<grid_for((long)n), 256, 0, s>>>(p, n, v);
}
void launch_fill_random(double* x, size_t n, unsigned long long seed, hipStream_t s) {
  if (n) fill_random_kernel<<<grid_for((long)std::min<size_t>(n, (size_t)1 << 40)), 256, 0, s>>>(x, n, seed);
}
void launch_fold_err(const int* err, int* info, int* sticky, hipStream_t s) {
  fold_err_kernel<<<1, 64, 0, s>>>(err, info, sticky);
}
void launch_fill_i32(int* p, size_t n, int v, hipStream_t s) {
  if (n) fill_kernel<int><<<grid_for((long)n), 256, 0, s>>>(p, n, v);
}

}  // namespace la
