// lu.cpp -- drivers of the LU solve path on one gfx950 GPU: DGETRF/SGETRF,
// DGETRS/SGETRS, DGESV, DLASWP, plus the per-device runtime (streams,
// workspace, tuning, error reporting).  Kernels live in the .hip files; this
// file sequences them.
//
// DGETRF (SRC/dgetrf.f:105-225) is run right-looking and blocked, exactly the
// reference's outer structure, but inside a transposed workspace T = A^T
// (A's rows are T's contiguous columns).  That turns every row interchange
// (DLASWP, SRC/dlaswp.f:156-166) into a contiguous, coalesced move, and the
// panel's rows into contiguous register loads.  Per block column j (width
// jb = NB):
//   panel     the DGETRF2 recursion (SRC/dgetrf2.f:220-263) flattened to
//             leaves of width <= 64, each factored by one cooperative kernel
//             (panel.hip), with the in-panel DLASWP / DTRSM / DGEMM between
//             leaves;
//   DLASWP    interchanges applied to the columns left and right of the
//             panel (SRC/dgetrf.f:193,199-200);
//   DTRSM     U12 = L11^-1 A12 (SRC/dgetrf.f:204-207) as a product with the
//             inverse of the unit triangle (aux.hip trtri + MFMA GEMM);
//   DGEMM     A22 -= A21*U12 (SRC/dgetrf.f:212-216) on the FP64 matrix cores.
// IPIV is the reference's 1-based interchange sequence with absolute rows;
// INFO the first exact zero pivot.
#include "la_common.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/lapack_amd.h"

namespace la {

// ---------------------------------------------------------------- runtime --

struct DevCtx {
  int dev = -1;
  hipStream_t stream = nullptr;  // stream of the synchronous entry points
  char* ws = nullptr;            // grow-only device workspace
  size_t ws_bytes = 0;
  char* io = nullptr;            // staging buffers for host-pointer calls
  size_t io_bytes = 0;
  hipEvent_t ws_free = nullptr;  // recorded after each use of ws
  bool ws_used = false;
  std::mutex mu;     // guards ws and the launch sequence of one call
  std::mutex io_mu;  // guards io across a whole synchronous call
};

static std::mutex g_mu;
static std::vector<DevCtx*> g_ctx;
static std::atomic<int> g_nb{0}, g_leaf{0};
static thread_local int t_last_error = 0;

static int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atoi(v);
}

int tune_nb() {
  int v = g_nb.load();
  if (v <= 0) v = env_int("LAPACK_AMD_NB", 256);
  return std::max(1, std::min(v, 256));  // the triangle-inverse kernel covers <= 256
}
int tune_leaf() {
  int v = g_leaf.load();
  if (v <= 0) v = env_int("LAPACK_AMD_LEAF", 64);
  return std::max(1, std::min(v, 64));
}

static int hip_check(hipError_t e) {
  if (e == hipSuccess) return 0;
  t_last_error = LA_ERR_HIP;
  std::fprintf(stderr, "lapack_amd: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
  return LA_ERR_HIP;
}
#define LA_TRY(expr)                           \
  do {                                         \
    int _rc = hip_check(expr);                 \
    if (_rc) return _rc;                       \
  } while (0)

static DevCtx* get_ctx() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
  std::lock_guard<std::mutex> g(g_mu);
  if ((int)g_ctx.size() <= dev) g_ctx.resize(dev + 1, nullptr);
  if (!g_ctx[dev]) {
    DevCtx* c = new DevCtx();
    c->dev = dev;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return nullptr; }
    if (hipEventCreateWithFlags(&c->ws_free, hipEventDisableTiming) != hipSuccess) { delete c; return nullptr; }
    g_ctx[dev] = c;
  }
  return g_ctx[dev];
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Reserve `bytes` of the device workspace for use on stream s.  Work from a
// previous call on another stream is ordered before it through ws_free.
static int ws_acquire(DevCtx* c, size_t bytes, hipStream_t s, char** out) {
  if (bytes > c->ws_bytes) {
    if (c->ws) {
      LA_TRY(hipDeviceSynchronize());
      LA_TRY(hipFree(c->ws));
      c->ws = nullptr;
      c->ws_bytes = 0;
    }
    size_t want = align_up(bytes, (size_t)1 << 21);
    if (hipMalloc(&c->ws, want) != hipSuccess) {
      (void)hipGetLastError();
      t_last_error = LA_ERR_NOMEM;
      return LA_ERR_NOMEM;
    }
    c->ws_bytes = want;
    c->ws_used = false;
  }
  if (c->ws_used) LA_TRY(hipStreamWaitEvent(s, c->ws_free, 0));
  *out = c->ws;
  return 0;
}
static int ws_release(DevCtx* c, hipStream_t s) {
  LA_TRY(hipEventRecord(c->ws_free, s));
  c->ws_used = true;
  return 0;
}

static int io_acquire(DevCtx* c, size_t bytes, char** out) {
  if (bytes > c->io_bytes) {
    if (c->io) { LA_TRY(hipStreamSynchronize(c->stream)); LA_TRY(hipFree(c->io)); c->io = nullptr; c->io_bytes = 0; }
    size_t want = align_up(bytes, (size_t)1 << 21);
    if (hipMalloc(&c->io, want) != hipSuccess) { (void)hipGetLastError(); t_last_error = LA_ERR_NOMEM; return LA_ERR_NOMEM; }
    c->io_bytes = want;
  }
  *out = c->io;
  return 0;
}

// Carves aligned sub-buffers out of one allocation.
struct Carve {
  size_t off = 0;
  template <class T> size_t take(size_t count) {
    off = align_up(off, 256);
    size_t o = off;
    off += sizeof(T) * std::max<size_t>(count, 1);
    return o;
  }
};

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// ---------------------------------------------------------- type plumbing --

template <class T> struct K;
template <> struct K<double> {
  static void gemm(char ta, char tb, int m, int n, int k, double al, const double* A, long lda, const double* B,
                   long ldb, double be, double* C, long ldc, hipStream_t s) {
    launch_dgemm(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc, s);
  }
  static void transpose(int r, int c, const double* in, long li, double* out, long lo, hipStream_t s) {
    launch_transpose_d(r, c, in, li, out, lo, s);
  }
  static void laswp_cols(double* T, long ldt, int r0, int r1, int k1, int k2, const int* ipiv, hipStream_t s) {
    launch_laswp_cols_d(T, ldt, r0, r1, k1, k2, ipiv, s);
  }
  static void gather(int nr, int nc, const int* perm, const double* src, long ls, double* dst, long ld, hipStream_t s) {
    launch_gather_rows_d(nr, nc, perm, src, ls, dst, ld, s);
  }
  static void trtri_blocks(bool up, bool unit, int n, int nb, const double* U, long ldu, double* W, long ldw,
                           hipStream_t s) {
    launch_trtri_blocks_d(up, unit, n, nb, U, ldu, W, ldw, s);
  }
  static void copy(int r, int c, const double* src, long ls, double* dst, long ld, hipStream_t s) {
    launch_copy_d(r, c, src, ls, dst, ld, s);
  }
  static void panel(double* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned* ctr, double* x,
                    int* err, hipStream_t s) {
    launch_getf2_panel_d(T, ldt, j0, w, m, ipiv, info, ctr, x, err, s);
  }
};
template <> struct K<float> {
  static void gemm(char ta, char tb, int m, int n, int k, float al, const float* A, long lda, const float* B,
                   long ldb, float be, float* C, long ldc, hipStream_t s) {
    launch_sgemm(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc, s);
  }
  static void transpose(int r, int c, const float* in, long li, float* out, long lo, hipStream_t s) {
    launch_transpose_s(r, c, in, li, out, lo, s);
  }
  static void laswp_cols(float* T, long ldt, int r0, int r1, int k1, int k2, const int* ipiv, hipStream_t s) {
    launch_laswp_cols_s(T, ldt, r0, r1, k1, k2, ipiv, s);
  }
  static void gather(int nr, int nc, const int* perm, const float* src, long ls, float* dst, long ld, hipStream_t s) {
    launch_gather_rows_s(nr, nc, perm, src, ls, dst, ld, s);
  }
  static void trtri_blocks(bool up, bool unit, int n, int nb, const float* U, long ldu, float* W, long ldw,
                           hipStream_t s) {
    launch_trtri_blocks_s(up, unit, n, nb, U, ldu, W, ldw, s);
  }
  static void copy(int r, int c, const float* src, long ls, float* dst, long ld, hipStream_t s) {
    launch_copy_s(r, c, src, ls, dst, ld, s);
  }
  static void panel(float* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned* ctr, double* x,
                    int* err, hipStream_t s) {
    launch_getf2_panel_s(T, ldt, j0, w, m, ipiv, info, ctr, x, err, s);
  }
};

// --------------------------------------------------------------- GETRF --

// Phase timing (bench/profiling): events around the phases of getrf.
struct PhaseTimer {
  bool on = false;
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> ev;
  std::vector<int> tag;
  void mark(int t) {
    if (!on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
    hipEventRecord(e, s);
    ev.push_back(e);
    tag.push_back(t);
  }
};
static std::atomic<int> g_phase_on{0};
static double g_phase_ms[LA_PHASE_COUNT];
static long g_phase_calls[LA_PHASE_COUNT];
static double g_phase_flops[LA_PHASE_COUNT];
static std::mutex g_phase_mu;

static void phase_collect(PhaseTimer& pt, const std::vector<double>& flops) {
  if (!pt.on) return;
  hipEventSynchronize(pt.ev.back());
  std::lock_guard<std::mutex> g(g_phase_mu);
  for (size_t i = 0; i + 1 < pt.ev.size(); ++i) {
    float ms = 0;
    hipEventElapsedTime(&ms, pt.ev[i], pt.ev[i + 1]);
    const int t = pt.tag[i];
    if (t >= 0 && t < LA_PHASE_COUNT) {
      g_phase_ms[t] += ms;
      g_phase_calls[t] += 1;
      g_phase_flops[t] += flops[i];
    }
  }
  for (auto e : pt.ev) hipEventDestroy(e);
  pt.ev.clear();
  pt.tag.clear();
}

template <class T>
size_t getrf_ws_bytes(int m, int n, int nb, int leaf) {
  const long ldt = (long)align_up((size_t)n, 32);
  Carve cv;
  cv.take<T>((size_t)ldt * m);                    // T workspace
  cv.take<T>((size_t)ldt * nb);                   // TRSM output
  cv.take<T>((size_t)nb * nb);                    // triangle inverse
  const int mn = std::min(m, n);
  cv.take<unsigned>((size_t)(mn / std::max(1, leaf) + mn / std::max(1, nb) + 8));  // barrier counters
  cv.take<char>(panel_xchg_bytes(0));
  cv.take<int>(4);                                // err flag
  return cv.off;
}

// Factor A (m x n, lda, device) in place; d_ipiv (min(m,n)) and d_info on
// device.  Asynchronous on s.
template <class T>
int getrf_device(int m, int n, T* A, long lda, int* ipiv, int* info, hipStream_t s) {
  DevCtx* c = get_ctx();
  if (!c) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(c->mu);
  if (m == 0 || n == 0) return 0;
  if (m > 65536) return LA_ERR_LIMIT;  // panel kernel: <= 128 workgroups x 512 rows
  const int nb = tune_nb();
  const int mn = std::min(m, n);
  const long ldt = (long)align_up((size_t)n, 32);
  const long ldx = ldt;
  char* ws = nullptr;
  int rc = ws_acquire(c, getrf_ws_bytes<T>(m, n, nb, tune_leaf()), s, &ws);
  if (rc) return rc;
  Carve cv;
  T* Tm = reinterpret_cast<T*>(ws + cv.take<T>((size_t)ldt * m));
  T* X = reinterpret_cast<T*>(ws + cv.take<T>((size_t)ldx * nb));
  T* W = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nb * nb));
  const size_t nctr = (size_t)(mn / std::max(1, tune_leaf()) + mn / nb + 8);
  unsigned* ctr = reinterpret_cast<unsigned*>(ws + cv.take<unsigned>(nctr));
  double* xchg = reinterpret_cast<double*>(ws + cv.take<char>(panel_xchg_bytes(0)));
  int* err = reinterpret_cast<int*>(ws + cv.take<int>(4));
  const long ldw = nb;

  PhaseTimer pt;
  pt.on = g_phase_on.load() != 0;
  pt.s = s;
  std::vector<double> fl;
  auto mark = [&](int tag, double flops) {
    if (!pt.on) return;
    pt.mark(tag);
    fl.push_back(flops);
  };

  launch_fill_u32(ctr, nctr, 0u, s);
  launch_fill_i32(err, 1, 0, s);
  launch_fill_i32(info, 1, 0, s);
  mark(LA_PHASE_TRANSPOSE, 0);
  K<T>::transpose(m, n, A, lda, Tm, ldt, s);  // T(c, r) = A(r, c)
  int leafno = 0;
  auto Tp = [&](long r, long col) { return Tm + r + col * ldt; };

  for (int j = 0; j < mn; j += nb) {
    const int jb = std::min(nb, mn - j);
    // ---- panel: A(j:m, j:j+jb) = T rows j..j+jb, T columns j..m
    for (int jj = j; jj < j + jb;) {
      const int h = m - jj;
      const int w = std::min({tune_leaf(), panel_max_width(h), j + jb - jj});
      mark(LA_PHASE_PANEL, 0);
      K<T>::panel(Tm, ldt, jj, w, m, ipiv, info, ctr + leafno++, xchg, err, s);
      mark(LA_PHASE_LASWP, 0);
      if (jj > j) K<T>::laswp_cols(Tm, ldt, j, jj, jj, jj + w, ipiv, s);
      const int rest = j + jb - jj - w;
      if (rest > 0) {
        K<T>::laswp_cols(Tm, ldt, jj + w, j + jb, jj, jj + w, ipiv, s);
        // in-panel DTRSM: T(jj+w:j+jb, jj:jj+w) <- * inv(U'(w x w))
        mark(LA_PHASE_TRSM, 2.0 * rest * w * w / 2.0);
        K<T>::trtri_blocks(true, true, w, w, Tp(jj, jj), ldt, W, ldw, s);
        K<T>::gemm('N', 'N', rest, w, w, T(1), Tp(jj + w, jj), ldt, W, ldw, T(0), X, ldx, s);
        K<T>::copy(rest, w, X, ldx, Tp(jj + w, jj), ldt, s);
        // in-panel DGEMM: T(jj+w:j+jb, jj+w:m) -= T(jj+w:j+jb, jj:jj+w) * T(jj:jj+w, jj+w:m)
        if (m - jj - w > 0) {
          mark(LA_PHASE_GEMM_PANEL, 2.0 * rest * (double)(m - jj - w) * w);
          K<T>::gemm('N', 'N', rest, m - jj - w, w, T(-1), Tp(jj + w, jj), ldt, Tp(jj, jj + w), ldt, T(1),
                     Tp(jj + w, jj + w), ldt, s);
        }
      }
      jj += w;
    }
    // ---- interchanges outside the panel (SRC/dgetrf.f:193,199-200)
    mark(LA_PHASE_LASWP, 0);
    if (j > 0) K<T>::laswp_cols(Tm, ldt, 0, j, j, j + jb, ipiv, s);
    if (j + jb < n) {
      K<T>::laswp_cols(Tm, ldt, j + jb, n, j, j + jb, ipiv, s);
      const int n2 = n - j - jb;
      // ---- DTRSM: U12^T = A12^T * inv(L11^T)
      mark(LA_PHASE_TRSM, 2.0 * n2 * jb * (double)jb / 2.0);
      K<T>::trtri_blocks(true, true, jb, jb, Tp(j, j), ldt, W, ldw, s);
      K<T>::gemm('N', 'N', n2, jb, jb, T(1), Tp(j + jb, j), ldt, W, ldw, T(0), X, ldx, s);
      K<T>::copy(n2, jb, X, ldx, Tp(j + jb, j), ldt, s);
      // ---- DGEMM: A22 -= A21*U12, i.e. T22 -= T12 * T21
      if (j + jb < m) {
        mark(LA_PHASE_GEMM, 2.0 * n2 * (double)(m - j - jb) * jb);
        K<T>::gemm('N', 'N', n2, m - j - jb, jb, T(-1), X, ldx, Tp(j, j + jb), ldt, T(1), Tp(j + jb, j + jb),
                   ldt, s);
      }
    }
  }
  mark(LA_PHASE_TRANSPOSE, 0);
  K<T>::transpose(n, m, Tm, ldt, A, lda, s);
  mark(-1, 0);
  launch_fold_err(err, info, s);
  rc = ws_release(c, s);
  phase_collect(pt, fl);
  return rc;
}

// --------------------------------------------------------------- GETRS --

template <class T>
size_t getrs_ws_bytes(int n, int nrhs, int nbs) {
  Carve cv;
  cv.take<int>((size_t)n);
  cv.take<T>((size_t)n * nrhs);
  cv.take<T>((size_t)nbs * nrhs);
  cv.take<T>((size_t)nbs * (size_t)((n + nbs - 1) / nbs) * nbs);
  cv.take<T>((size_t)nbs * (size_t)((n + nbs - 1) / nbs) * nbs);
  return cv.off;
}

// Solve op(A) X = B with the LU of getrf (SRC/dgetrs.f:181-218).  Device
// pointers, asynchronous on s.  Blocked substitution: the diagonal blocks
// of L and U are inverted once, every block step is then two MFMA GEMMs.
template <class T>
int getrs_device(char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb,
                 hipStream_t s) {
  DevCtx* c = get_ctx();
  if (!c) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(c->mu);
  if (n == 0 || nrhs == 0) return 0;
  const bool notran = (trans == 'N' || trans == 'n');
  const int nbs = 256;
  const int nblk = (n + nbs - 1) / nbs;
  char* ws = nullptr;
  int rc = ws_acquire(c, getrs_ws_bytes<T>(n, nrhs, nbs), s, &ws);
  if (rc) return rc;
  Carve cv;
  int* perm = reinterpret_cast<int*>(ws + cv.take<int>((size_t)n));
  T* X = reinterpret_cast<T*>(ws + cv.take<T>((size_t)n * nrhs));
  T* Y = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nbs * nrhs));
  T* WL = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nbs * nblk * nbs));
  T* WU = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nbs * nblk * nbs));
  const long ldx = n, ldy = nbs, ldw = nbs;
  auto Ap = [&](long r, long col) { return A + r + col * lda; };

  PhaseTimer pt;
  pt.on = g_phase_on.load() != 0;
  pt.s = s;
  std::vector<double> fl;
  pt.mark(LA_PHASE_GETRS);
  fl.push_back(2.0 * (double)n * n * nrhs);

  K<T>::trtri_blocks(false, true, n, nbs, A, lda, WL, ldw, s);   // inv(L_dd), unit lower
  K<T>::trtri_blocks(true, false, n, nbs, A, lda, WU, ldw, s);   // inv(U_dd), non-unit upper
  if (notran) {
    launch_swaps_to_perm(n, 0, n, ipiv, 1, perm, s);
    K<T>::gather(n, nrhs, perm, B, ldb, X, ldx, s);                // X = P*B (dgetrs.f:187)
    for (int d = 0; d < nblk; ++d) {                               // L Y = X (dgetrs.f:191)
      const int b0 = d * nbs, bs = std::min(nbs, n - b0);
      K<T>::gemm('N', 'N', bs, nrhs, bs, T(1), WL + (long)d * nbs * ldw, ldw, X + b0, ldx, T(0), Y, ldy, s);
      K<T>::copy(bs, nrhs, Y, ldy, X + b0, ldx, s);
      if (b0 + bs < n)
        K<T>::gemm('N', 'N', n - b0 - bs, nrhs, bs, T(-1), Ap(b0 + bs, b0), lda, X + b0, ldx, T(1), X + b0 + bs,
                   ldx, s);
    }
    for (int d = nblk - 1; d >= 0; --d) {                          // U X = Y (dgetrs.f:197)
      const int b0 = d * nbs, bs = std::min(nbs, n - b0);
      K<T>::gemm('N', 'N', bs, nrhs, bs, T(1), WU + (long)d * nbs * ldw, ldw, X + b0, ldx, T(0), Y, ldy, s);
      K<T>::copy(bs, nrhs, Y, ldy, X + b0, ldx, s);
      if (b0 > 0)
        K<T>::gemm('N', 'N', b0, nrhs, bs, T(-1), Ap(0, b0), lda, X + b0, ldx, T(1), X, ldx, s);
    }
    K<T>::copy(n, nrhs, X, ldx, B, ldb, s);
  } else {
    K<T>::copy(n, nrhs, B, ldb, X, ldx, s);
    for (int d = 0; d < nblk; ++d) {                               // U^T Y = B (dgetrs.f:205)
      const int b0 = d * nbs, bs = std::min(nbs, n - b0);
      K<T>::gemm('T', 'N', bs, nrhs, bs, T(1), WU + (long)d * nbs * ldw, ldw, X + b0, ldx, T(0), Y, ldy, s);
      K<T>::copy(bs, nrhs, Y, ldy, X + b0, ldx, s);
      if (b0 + bs < n)
        K<T>::gemm('T', 'N', n - b0 - bs, nrhs, bs, T(-1), Ap(b0, b0 + bs), lda, X + b0, ldx, T(1), X + b0 + bs,
                   ldx, s);
    }
    for (int d = nblk - 1; d >= 0; --d) {                          // L^T X = Y (dgetrs.f:211)
      const int b0 = d * nbs, bs = std::min(nbs, n - b0);
      K<T>::gemm('T', 'N', bs, nrhs, bs, T(1), WL + (long)d * nbs * ldw, ldw, X + b0, ldx, T(0), Y, ldy, s);
      K<T>::copy(bs, nrhs, Y, ldy, X + b0, ldx, s);
      if (b0 > 0)
        K<T>::gemm('T', 'N', b0, nrhs, bs, T(-1), Ap(b0, 0), lda, X + b0, ldx, T(1), X, ldx, s);
    }
    launch_swaps_to_perm(n, 0, n, ipiv, -1, perm, s);              // dgetrs.f:217
    K<T>::gather(n, nrhs, perm, X, ldx, B, ldb, s);
  }
  pt.mark(-1);
  fl.push_back(0);
  rc = ws_release(c, s);
  phase_collect(pt, fl);
  return rc;
}

// DLASWP on a column-major device matrix (SRC/dlaswp.f:112-189).  nrows is
// the number of rows the interchanges may reach.
template <class T>
int laswp_device(int ncols, T* A, long lda, int nrows, int k1, int k2, const int* ipiv, int incx,
                 hipStream_t s) {
  DevCtx* c = get_ctx();
  if (!c) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(c->mu);
  if (ncols <= 0 || incx == 0 || k2 < k1 || nrows <= 0) return 0;
  char* ws = nullptr;
  Carve cv0;
  cv0.take<int>((size_t)nrows);
  cv0.take<T>((size_t)nrows * ncols);
  int rc = ws_acquire(c, cv0.off, s, &ws);
  if (rc) return rc;
  Carve cv;
  int* perm = reinterpret_cast<int*>(ws + cv.take<int>((size_t)nrows));
  T* X = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nrows * ncols));
  // k1, k2 are 1-based and inclusive in the reference; perm kernel takes
  // 0-based [k1-1, k2).
  launch_swaps_to_perm(nrows, k1 - 1, k2, ipiv, incx, perm, s);
  K<T>::gather(nrows, ncols, perm, A, lda, X, nrows, s);
  K<T>::copy(nrows, ncols, X, nrows, A, lda, s);
  return ws_release(c, s);
}

// Explicit instantiations used by the C ABI and the mixed-precision solver.
template int getrf_device<double>(int, int, double*, long, int*, int*, hipStream_t);
template int getrf_device<float>(int, int, float*, long, int*, int*, hipStream_t);
template int getrs_device<double>(char, int, int, const double*, long, const int*, double*, long, hipStream_t);
template int getrs_device<float>(char, int, int, const float*, long, const int*, float*, long, hipStream_t);
template int laswp_device<double>(int, double*, long, int, int, int, const int*, int, hipStream_t);
template int laswp_device<float>(int, float*, long, int, int, int, const int*, int, hipStream_t);

int last_error() { return t_last_error; }
void set_tuning(int nb, int leaf) { g_nb = nb; g_leaf = leaf; }
DevCtx* ctx_for_current() { return get_ctx(); }
hipStream_t ctx_stream(DevCtx* c) { return c->stream; }
int io_buffer(DevCtx* c, size_t bytes, char** out) { return io_acquire(c, bytes, out); }
std::mutex& io_mutex(DevCtx* c) { return c->io_mu; }
void phase_enable(int on) { g_phase_on = on; }
void phase_read(double* ms, long* calls, double* flops) {
  std::lock_guard<std::mutex> g(g_phase_mu);
  for (int i = 0; i < LA_PHASE_COUNT; ++i) {
    if (ms) ms[i] = g_phase_ms[i];
    if (calls) calls[i] = g_phase_calls[i];
    if (flops) flops[i] = g_phase_flops[i];
  }
}
void phase_reset() {
  std::lock_guard<std::mutex> g(g_phase_mu);
  for (int i = 0; i < LA_PHASE_COUNT; ++i) { g_phase_ms[i] = 0; g_phase_calls[i] = 0; g_phase_flops[i] = 0; }
}

}  // namespace la
