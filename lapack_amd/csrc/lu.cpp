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
#include <cctype>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
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
  hipStream_t pstream = nullptr;      // panel stream (look-ahead)
  hipStream_t ustream = nullptr;      // trailing-update stream
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  char* mp = nullptr;            // buffers of the mixed-precision solvers
  size_t mp_bytes = 0;
  int* status = nullptr;         // sticky device word: bit 0 = a bounded wait gave up
  bool xcd8 = false;             // 8 XCDs of >= 32 CUs: the panel may own XCD 0
  // flag base of the next panel launch: grows by 64 per launch across calls,
  // so that an exchange word left from an earlier launch (in memory or in an
  // XCD's L2) never carries a current flag
  unsigned leaf_tag = 0;
  std::mutex mu;     // guards ws and the launch sequence of one call
  std::mutex io_mu;  // guards io across a whole synchronous call
  std::mutex mp_mu;  // guards mp across a whole mixed-precision solve
};

static std::mutex g_mu;
static std::vector<DevCtx*> g_ctx;
static std::atomic<int> g_nb{0}, g_leaf{0};
static thread_local int t_last_error = 0;
// Set while a synchronous entry point redoes a call whose bounded wait gave
// up: every leaf then takes the panel form without waits and DGETRS the
// blocked form without waits.
static thread_local int t_safe = 0;
// Test hook: the next g_fail_waits factorizations / solves start with their
// wait-error word already set, so every bounded wait in them gives up at
// once (each kernel still runs to its end).
static std::atomic<int> g_fail_waits{0};

static int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atoi(v);
}

// ---- block size NB and panel leaf width ---------------------------------------
// In order: lapack_amd_set_tuning (the XLAENV analogue), the environment
// (LAPACK_AMD_NB, LAPACK_AMD_LEAF), the tuned table for the order's bucket
// (lapack_amd_tune_getrf / lapack_amd_load_tuning, or the file named by
// LAPACK_AMD_TUNE_FILE, read once), then the defaults 256 / 64.
struct TuneEntry {
  int nb = 0, leaf = 0;
  double ms = 0.0;
};
static std::mutex g_tune_mu;
static std::map<int, TuneEntry> g_tune;  // bucket -> entry
static bool g_tune_file_read = false;

int tune_bucket(int n) {
  int b = 256;
  while (b < n && b < (1 << 30)) b *= 2;
  return b;
}
static int load_tuning_locked(const char* path);
static TuneEntry table_entry(int n) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  if (!g_tune_file_read) {
    g_tune_file_read = true;
    const char* f = std::getenv("LAPACK_AMD_TUNE_FILE");
    if (f && *f) (void)load_tuning_locked(f);
  }
  auto it = g_tune.find(tune_bucket(n));
  return it == g_tune.end() ? TuneEntry{} : it->second;
}
static thread_local int t_force_nb = 0, t_force_leaf = 0;  // a tuning trial's candidate
int tune_nb_for(int n) {
  if (t_force_nb > 0) return t_force_nb;
  int v = g_nb.load();
  if (v <= 0) v = env_int("LAPACK_AMD_NB", 0);
  if (v <= 0 && n > 0) v = table_entry(n).nb;
  if (v <= 0) v = 256;
  return std::max(1, std::min(v, 256));  // the triangle-inverse kernel covers <= 256
}
int tune_leaf_for(int n) {
  if (t_force_leaf > 0) return t_force_leaf;
  int v = g_leaf.load();
  if (v <= 0) v = env_int("LAPACK_AMD_LEAF", 0);
  if (v <= 0 && n > 0) v = table_entry(n).leaf;
  if (v <= 0) v = 64;
  return std::max(1, std::min(v, 64));
}
int tune_nb() { return tune_nb_for(0); }
int tune_leaf() { return tune_leaf_for(0); }

// The table as JSON: {"getrf": {"<bucket>": {"nb": .., "leaf": .., "ms": ..}, ..}}
static int load_tuning_locked(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return -1;
  std::string txt;
  char buf[4096];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, k);
  std::fclose(f);
  // every "<bucket>": {"nb": .., "leaf": .., "ms": ..} after the "getrf" key
  size_t pos = txt.find("\"getrf\"");
  int count = 0;
  while (pos != std::string::npos) {
    const size_t q = txt.find('"', pos + 1 + (count == 0 ? 6 : 0));
    if (q == std::string::npos || q + 1 >= txt.size() || !std::isdigit((unsigned char)txt[q + 1])) break;
    const int bucket = std::atoi(txt.c_str() + q + 1);
    const size_t open = txt.find('{', q), close = txt.find('}', q);
    if (open == std::string::npos || close == std::string::npos || open > close) break;
    const std::string body = txt.substr(open, close - open);
    auto field = [&](const char* key, double dflt) {
      const size_t k = body.find(key);
      return k == std::string::npos ? dflt : std::atof(body.c_str() + k + std::strlen(key));
    };
    TuneEntry e;
    e.nb = (int)field("\"nb\":", 0);
    e.leaf = (int)field("\"leaf\":", 0);
    e.ms = field("\"ms\":", 0.0);
    if (bucket > 0 && e.nb > 0 && e.leaf > 0) {
      g_tune[bucket] = e;
      ++count;
    }
    pos = close;
  }
  return count;
}
static int save_tuning_locked(const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  std::fprintf(f, "{\"getrf\": {");
  bool first = true;
  for (auto& kv : g_tune) {
    std::fprintf(f, "%s\"%d\": {\"nb\": %d, \"leaf\": %d, \"ms\": %.4f}", first ? "" : ", ", kv.first,
                 kv.second.nb, kv.second.leaf, kv.second.ms);
    first = false;
  }
  std::fprintf(f, "}}\n");
  std::fclose(f);
  return 0;
}
int load_tuning(const char* path) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune_file_read = true;
  if (!path) {
    g_tune.clear();
    return 0;
  }
  return load_tuning_locked(path);
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
    if (hipMalloc(&c->status, 256) != hipSuccess || hipMemset(c->status, 0, 256) != hipSuccess) {
      (void)hipGetLastError();
      delete c;
      return nullptr;
    }
    int nxcc = 0, ncu = 0;
    if (hipDeviceGetAttribute(&nxcc, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      (void)hipGetLastError();
      nxcc = 0;
    }
    c->xcd8 = nxcc == 8 && ncu >= 8 * 32;
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

static int mp_acquire(DevCtx* c, size_t bytes, char** out) {
  if (bytes > c->mp_bytes) {
    if (c->mp) { LA_TRY(hipDeviceSynchronize()); LA_TRY(hipFree(c->mp)); c->mp = nullptr; c->mp_bytes = 0; }
    size_t want = align_up(bytes, (size_t)1 << 21);
    if (hipMalloc(&c->mp, want) != hipSuccess) { (void)hipGetLastError(); t_last_error = LA_ERR_NOMEM; return LA_ERR_NOMEM; }
    c->mp_bytes = want;
  }
  *out = c->mp;
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
                   long ldb, double be, double* C, long ldc, hipStream_t s, int nskip = 0) {
    launch_dgemm(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc, s, nskip);
  }
  static void update(int m, int n, int k, const double* A, long lda, const double* B, long ldb, double* C, long ldc,
                     int nskip, hipStream_t s) {
    launch_dgemm_update(m, n, k, A, lda, B, ldb, C, ldc, nskip, s);
  }
  static void transpose(int r, int c, const double* in, long li, double* out, long lo, hipStream_t s,
                        const int* skip = nullptr) {
    launch_transpose_d(r, c, in, li, out, lo, s, skip);
  }
  static void apply_cols(double* T, long ldt, int r0, int r1, int r2, int r3, const int* pairs, const int* np,
                         hipStream_t s, int nskip = 0) {
    launch_apply_pairs_cols_d(T, ldt, r0, r1, r2, r3, pairs, np, s, nskip);
  }
  static void gather(int nr, int nc, const int* perm, const double* src, long ls, double* dst, long ld, hipStream_t s,
                     const int* skip = nullptr) {
    launch_gather_rows_d(nr, nc, perm, src, ls, dst, ld, s, skip);
  }
  static void permute_rows(int nr, int nc, const int* perm, double* A, long lda, hipStream_t s) {
    launch_permute_rows_d(nr, nc, perm, A, lda, s);
  }
  static void trtri_blocks(bool up, bool unit, int n, int nb, const double* U, long ldu, double* W, long ldw,
                           hipStream_t s) {
    launch_trtri_blocks_d(up, unit, n, nb, U, ldu, W, ldw, s);
  }
  static void copy(int r, int c, const double* src, long ls, double* dst, long ld, hipStream_t s, int nskip = 0,
                   const int* skip = nullptr) {
    launch_copy_d(r, c, src, ls, dst, ld, s, nskip, skip);
  }
  static int panel(double* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase, double* x,
                   int* err, int mode, hipStream_t s) {
    return launch_getf2_panel_d(T, ldt, j0, w, m, ipiv, info, tagbase, x, err, mode, s);
  }
  static void boundary(double* T, long ldt, int j, int jj, int w, int jend, int m, const int* ipiv, hipStream_t s) {
    launch_panel_boundary_d(T, ldt, j, jj, w, jend, m, ipiv, s);
  }
};
// Operand precision of the FP32 factorization's trailing updates on the
// calling thread: 0 = FP32, 1 = FP16, 2 = BF16 (set by getrf_device_lp for
// the duration of one factorization).
static thread_local int t_upd_kind = 0;

template <> struct K<float> {
  static void gemm(char ta, char tb, int m, int n, int k, float al, const float* A, long lda, const float* B,
                   long ldb, float be, float* C, long ldc, hipStream_t s, int nskip = 0) {
    launch_sgemm(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc, s, nskip);
  }
  static void update(int m, int n, int k, const float* A, long lda, const float* B, long ldb, float* C, long ldc,
                     int nskip, hipStream_t s) {
    if (t_upd_kind == 1 || t_upd_kind == 2)
      launch_sgemm_update_half(t_upd_kind, m, n, k, A, lda, B, ldb, C, ldc, nskip, s);
    else
      launch_sgemm_update(m, n, k, A, lda, B, ldb, C, ldc, nskip, s);
  }
  static void transpose(int r, int c, const float* in, long li, float* out, long lo, hipStream_t s,
                        const int* skip = nullptr) {
    launch_transpose_s(r, c, in, li, out, lo, s, skip);
  }
  static void apply_cols(float* T, long ldt, int r0, int r1, int r2, int r3, const int* pairs, const int* np,
                         hipStream_t s, int nskip = 0) {
    launch_apply_pairs_cols_s(T, ldt, r0, r1, r2, r3, pairs, np, s, nskip);
  }
  static void gather(int nr, int nc, const int* perm, const float* src, long ls, float* dst, long ld, hipStream_t s,
                     const int* skip = nullptr) {
    launch_gather_rows_s(nr, nc, perm, src, ls, dst, ld, s, skip);
  }
  static void permute_rows(int nr, int nc, const int* perm, float* A, long lda, hipStream_t s) {
    launch_permute_rows_s(nr, nc, perm, A, lda, s);
  }
  static void trtri_blocks(bool up, bool unit, int n, int nb, const float* U, long ldu, float* W, long ldw,
                           hipStream_t s) {
    launch_trtri_blocks_s(up, unit, n, nb, U, ldu, W, ldw, s);
  }
  static void copy(int r, int c, const float* src, long ls, float* dst, long ld, hipStream_t s, int nskip = 0,
                   const int* skip = nullptr) {
    launch_copy_s(r, c, src, ls, dst, ld, s, nskip, skip);
  }
  static int panel(float* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase, double* x,
                   int* err, int mode, hipStream_t s) {
    return launch_getf2_panel_s(T, ldt, j0, w, m, ipiv, info, tagbase, x, err, mode, s);
  }
  static void boundary(float* T, long ldt, int j, int jj, int w, int jend, int m, const int* ipiv, hipStream_t s) {
    launch_panel_boundary_s(T, ldt, j, jj, w, jend, m, ipiv, s);
  }
};

// --------------------------------------------------------------- GETRF --

// Phase timing (bench/profiling): events around the phases of getrf.
// Level 1 times every phase; level 2 only the trailing update (two events
// per update launch), which keeps the events' own cost out of a timed run.
struct PhaseTimer {
  bool on = false;
  int level = 1;
  bool skipping = false;  // the open interval is not recorded
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> ev;
  std::vector<int> tag;
  // t >= 0 opens an interval of phase t, t < 0 closes it; true if recorded
  bool mark(int t) {
    if (!on) return false;
    if (t >= 0) {
      skipping = level == 2 && t != LA_PHASE_GEMM;
      if (skipping) return false;
    } else if (skipping) {
      skipping = false;
      return false;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { on = false; return false; }
    hipEventRecord(e, s);
    ev.push_back(e);
    tag.push_back(t);
    return true;
  }
};
static std::atomic<int> g_phase_on{0};
static double g_phase_ms[LA_PHASE_COUNT];
static long g_phase_calls[LA_PHASE_COUNT];
static double g_phase_flops[LA_PHASE_COUNT];
static std::mutex g_phase_mu;
struct PendingPhases {
  std::vector<hipEvent_t> ev;
  std::vector<int> tag;
  std::vector<double> flops;
};
static std::vector<PendingPhases> g_pending;

// Hand the call's events over; they are resolved (synchronising on them)
// only when the phase totals are read, so timing adds no host sync.
static void phase_collect(PhaseTimer& pt, const std::vector<double>& flops) {
  if (!pt.on) return;
  std::lock_guard<std::mutex> g(g_phase_mu);
  g_pending.push_back(PendingPhases{pt.ev, pt.tag, flops});
  pt.ev.clear();
  pt.tag.clear();
}
static void phase_resolve_locked() {
  for (auto& pp : g_pending) {
    if (pp.ev.empty()) continue;
    hipEventSynchronize(pp.ev.back());
    for (size_t i = 0; i + 1 < pp.ev.size(); ++i) {
      float ms = 0;
      hipEventElapsedTime(&ms, pp.ev[i], pp.ev[i + 1]);
      const int t = pp.tag[i];
      if (t >= 0 && t < LA_PHASE_COUNT) {
        g_phase_ms[t] += ms;
        g_phase_calls[t] += 1;
        g_phase_flops[t] += pp.flops[i];
      }
    }
    for (auto e : pp.ev) hipEventDestroy(e);
  }
  g_pending.clear();
}

template <class T>
size_t getrf_ws_bytes(int m, int n, int nb, int leaf) {
  (void)leaf;
  const long ldt = (long)align_up((size_t)n, 32);
  Carve cv;
  cv.take<T>((size_t)ldt * m);                    // T workspace
  for (int i = 0; i < 2; ++i) {                   // per stream: TRSM output, triangle inverse, pairs
    cv.take<T>((size_t)ldt * nb);
    cv.take<T>((size_t)nb * nb);
    cv.take<int>((size_t)2 * swp_max_pairs() + 64);
  }
  cv.take<char>(panel_xchg_bytes(0));             // panel exchange + tall-leaf scratch
  cv.take<int>(4);                                // err flag
  cv.take<int>((size_t)m);                        // position map (fallback)
  return cv.off;
}

// Factor A (m x n, lda, device) in place; d_ipiv (min(m,n)) and d_info on
// device.  Asynchronous on s.  With look-ahead: the trailing update of block
// j first updates the columns of panel j+1, which is then factored on the
// device's panel stream while the rest of the update runs on s.
template <class T>
int getrf_device(int m, int n, T* A, long lda, int* ipiv, int* info, hipStream_t s) {
  DevCtx* c = get_ctx();
  if (!c) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(c->mu);
  if (m == 0 || n == 0) {
    launch_fill_i32(info, 1, 0, s);  // INFO = 0 as the reference's quick return
    return 0;
  }
  const int mn = std::min(m, n);
  const int nb = tune_nb_for(mn);
  const int leaf = tune_leaf_for(mn);
  const long ldt = (long)align_up((size_t)n, 32);
  const long ldx = ldt;
  char* ws = nullptr;
  int rc = ws_acquire(c, getrf_ws_bytes<T>(m, n, nb, leaf), s, &ws);
  if (rc) return rc;
  if (!c->pstream) {
    // the panel chain (latency-bound) and the trailing update run on two
    // streams of their own, fenced to the caller's stream
    LA_TRY(hipStreamCreateWithFlags(&c->pstream, hipStreamNonBlocking));
    LA_TRY(hipStreamCreateWithFlags(&c->ustream, hipStreamNonBlocking));
    for (int i = 0; i < 5; ++i) LA_TRY(hipEventCreateWithFlags(&c->ev[i], hipEventDisableTiming));
  }
  hipStream_t sp = c->pstream;
  hipStream_t s_user = s;
  // all work of this call runs on the library's two streams, fenced to the
  // caller's stream at entry and exit
  LA_TRY(hipEventRecord(c->ev[2], s_user));
  s = c->ustream;
  LA_TRY(hipStreamWaitEvent(s, c->ev[2], 0));
  Carve cv;
  T* Tm = reinterpret_cast<T*>(ws + cv.take<T>((size_t)ldt * m));
  T* Xs[2];
  T* Ws[2];
  int* Ps[2];
  for (int i = 0; i < 2; ++i) {
    Xs[i] = reinterpret_cast<T*>(ws + cv.take<T>((size_t)ldx * nb));
    Ws[i] = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nb * nb));
    Ps[i] = reinterpret_cast<int*>(ws + cv.take<int>((size_t)2 * swp_max_pairs() + 64));
  }
  double* xchg = reinterpret_cast<double*>(ws + cv.take<char>(panel_xchg_bytes(0)));
  int* err = reinterpret_cast<int*>(ws + cv.take<int>(4));
  int* gmap = reinterpret_cast<int*>(ws + cv.take<int>((size_t)m));
  const long ldw = nb;
  // the multi-workgroup panel owns XCD 0 only on a device of 8 XCDs; a
  // redone call (t_safe) takes the panel form without waits
  const bool xcdsplit = c->xcd8 && env_int("LAPACK_AMD_PANEL_XCD", 1) != 0 && !t_safe;
  int fail = g_fail_waits.load();
  while (fail > 0 && !g_fail_waits.compare_exchange_weak(fail, fail - 1)) {
  }
  // panel workgroups from which the rest's interchanges and DTRSM go first
  const int pre_pg = env_int("LAPACK_AMD_PRE_PG", 27);

  PhaseTimer pa, pp;  // main stream, panel stream
  pa.on = pp.on = g_phase_on.load() != 0;
  pa.level = pp.level = g_phase_on.load();
  pa.s = s;
  pp.s = sp;
  std::vector<double> fla, flp;
  // Phase intervals: begin(tag) ... end() around the launches of one phase,
  // so time spent waiting on the other stream is not attributed to it.
  auto mark = [&](PhaseTimer& pt, std::vector<double>& fl, int tag, double flops) {
    if (pt.mark(tag)) fl.push_back(flops);
  };
  auto endm = [&](PhaseTimer& pt, std::vector<double>& fl) { mark(pt, fl, -1, 0); };
  auto Tp = [&](long r, long col) { return Tm + r + col * ldt; };

  // the panel's flag-in-data exchange needs a zeroed buffer per factorization;
  // the flag base starts over only long before it could wrap
  if (c->leaf_tag > 0xF0000000u) c->leaf_tag = 0;
  launch_fill_u32(reinterpret_cast<unsigned*>(xchg), panel_xchg_bytes(0) / sizeof(unsigned), 0u, s);
  launch_fill_i32(err, 1, fail > 0 ? 1 : 0, s);
  launch_fill_i32(info, 1, 0, s);
  mark(pa, fla, LA_PHASE_TRANSPOSE, 0);
  K<T>::transpose(m, n, A, lda, Tm, ldt, s);  // T(c, r) = A(r, c)
  endm(pa, fla);

  // Panel j: A(j:m, j:j+jb) = T rows j..j+jb, T columns j..m; the DGETRF2
  // recursion flattened to leaves, with in-panel interchanges, DTRSM and
  // DGEMM between leaves (buffers of slot 1, stream st).
  auto factor_panel = [&](int j, int jb, hipStream_t st, PhaseTimer& pt, std::vector<double>& fl,
                          bool xcd0) -> int {
    const int mode = t_safe ? PANEL_TALL : xcd0 ? PANEL_XCD0 : PANEL_PLAIN;
    for (int jj = j; jj < j + jb;) {
      const int h = m - jj;
      const int w = std::min({leaf, panel_max_width(h), j + jb - jj});
      mark(pt, fl, LA_PHASE_PANEL, 0);
      c->leaf_tag += 64u;
      int r = K<T>::panel(Tm, ldt, jj, w, m, ipiv, info, c->leaf_tag, xchg, err, mode, st);
      endm(pt, fl);
      if (r) return r;
      const int rest = j + jb - jj - w;
      // interchanges on the panel's other columns and the in-panel DTRSM
      // (SRC/dgetrf2.f:234-241): one single-workgroup launch
      if (jj > j || rest > 0) {
        mark(pt, fl, LA_PHASE_LASWP, 0);
        K<T>::boundary(Tm, ldt, j, jj, w, j + jb, m, ipiv, st);
        endm(pt, fl);
      }
      // in-panel DGEMM: T(jj+w:j+jb, jj+w:m) -= T(jj+w:j+jb, jj:jj+w) * T(jj:jj+w, jj+w:m)
      if (rest > 0 && m - jj - w > 0) {
        mark(pt, fl, LA_PHASE_GEMM_PANEL, 2.0 * rest * (double)(m - jj - w) * w);
        K<T>::gemm('N', 'N', rest, m - jj - w, w, T(-1), Tp(jj + w, jj), ldt, Tp(jj, jj + w), ldt, T(1),
                   Tp(jj + w, jj + w), ldt, st);
        endm(pt, fl);
      }
      jj += w;
    }
    return 0;
  };

  // W = inv(L11^T) of panel jp (the DTRSM's triangle) into its buffer;
  // buffers alternate between consecutive panels
  auto inv_tri = [&](int jp, int jbp, hipStream_t st, PhaseTimer& pt, std::vector<double>& fl) {
    mark(pt, fl, LA_PHASE_TRSM, 0);
    K<T>::trtri_blocks(true, true, jbp, jbp, Tp(jp, jp), ldt, Ws[(jp / nb) & 1], ldw, st);
    endm(pt, fl);
  };
  // panel 0 on the main stream
  rc = factor_panel(0, std::min(nb, mn), s, pa, fla, false);
  if (rc == 0 && n > std::min(nb, mn)) inv_tri(0, std::min(nb, mn), s, pa, fla);
  for (int j = 0; j < mn && rc == 0; j += nb) {
    const int jb = std::min(nb, mn - j);
    if (j > 0) LA_TRY(hipStreamWaitEvent(s, c->ev[0], 0));  // panel j done (panel stream)
    T* X = Xs[0];
    T* W = Ws[(j / nb) & 1];
    int* pairs = Ps[0];
    int* npairs = pairs + 2 * swp_max_pairs();
    // Only the next panel's columns are on the critical path: their
    // interchanges, TRSM and update run first and release the panel stream;
    // the interchanges left of the panel and the TRSM / update of the
    // columns right of the next panel follow on s, beside that panel.
    const int jn = j + jb;
    const int n2 = n - jn;
    const int jbn = jn < mn ? std::min(nb, mn - jn) : 0;
    const int mrows = m - jn;
    // the next panel takes XCD 0 for itself (32 workgroups, one per CU) when
    // it runs multi-workgroup; the rest of the step leaves it.  A taller
    // panel (more groups) shares the GPU instead: measured at n = 32768,
    // reserving a second XCD slows the then dominant update more than it
    // speeds the panel (in FP32 as well: DSGESV 467 instead of 406 ms).
    const int pg = jbn > 0 ? panel_groups(m - jn) : 0;
    const int npx = (xcdsplit && pg > 0 && pg <= 32) ? 1 : 0;
    const bool split = npx > 0;
    // While the panel holds (nearly) all of XCD 0, every launch beside it
    // waits for slots there even for the workgroups that only leave: in the
    // first steps (measured: 28 or more panel workgroups) the rest-of-matrix
    // DTRSM product then took 5x its time.  There the rest's interchanges,
    // DTRSM and copy run before the panel, on the whole GPU (those steps are
    // bound by the trailing update, not by the panel chain).
    const bool pre = split && pg >= pre_pg;
    // ---- interchanges outside the panel (SRC/dgetrf.f:193,199-200): next panel's columns
    mark(pa, fla, LA_PHASE_LASWP, 0);
    if (j > 0 || jn < n) launch_swaps_to_pairs(j, jn, m, ipiv, pairs, npairs, gmap, s);
    if (jbn > 0) K<T>::apply_cols(Tm, ldt, jn, jn + jbn, 0, 0, pairs, npairs, s);
    endm(pa, fla);
    // the triangle's inverse, computed on the panel stream after panel j
    if (j > 0 && n2 > 0) LA_TRY(hipStreamWaitEvent(s, c->ev[4], 0));
    if (jbn > 0) {
      // ---- DTRSM: U12^T = A12^T * inv(L11^T)  (SRC/dgetrf.f:204-207), next panel's columns
      //      into X (copied back into T after that panel has been released)
      mark(pa, fla, LA_PHASE_TRSM, 2.0 * jbn * jb * (double)jb / 2.0);
      K<T>::gemm('N', 'N', jbn, jb, jb, T(1), Tp(jn, j), ldt, W, ldw, T(0), X, ldx, s);
      endm(pa, fla);
    }
    // ---- DGEMM: A22 -= A21*U12, i.e. T22 -= T12 * T21 (SRC/dgetrf.f:212-216),
    //      the next panel's columns, then that panel on the panel stream
    const int nr = n2 - jbn;
    // ---- the rest's interchanges (left of this panel, right of the next one)
    //      and DTRSM: here, before the next panel, or beside it below
    auto rest_swap_trsm = [&](int nsk) {
      mark(pa, fla, LA_PHASE_LASWP, 0);
      if (j > 0 || jn + jbn < n) K<T>::apply_cols(Tm, ldt, 0, j, jn + jbn, n, pairs, npairs, s, nsk);
      endm(pa, fla);
      if (nr > 0) {
        mark(pa, fla, LA_PHASE_TRSM, 2.0 * nr * jb * (double)jb / 2.0);
        K<T>::gemm('N', 'N', nr, jb, jb, T(1), Tp(jn + jbn, j), ldt, W, ldw, T(0), X + jbn, ldx, s, nsk);
        K<T>::copy(nr, jb, X + jbn, ldx, Tp(jn + jbn, j), ldt, s, nsk);
        endm(pa, fla);
      }
    };
    if (jbn > 0 && mrows > 0) {
      mark(pa, fla, LA_PHASE_GEMM, 2.0 * jbn * (double)mrows * jb);
      K<T>::update(jbn, mrows, jb, X, ldx, Tp(j, jn), ldt, Tp(jn, jn), ldt, false, s);
      endm(pa, fla);
      if (pre) rest_swap_trsm(0);
      LA_TRY(hipEventRecord(c->ev[1], s));
      LA_TRY(hipStreamWaitEvent(sp, c->ev[1], 0));
      rc = factor_panel(jn, jbn, sp, pp, flp, split);
      if (rc) break;
      LA_TRY(hipEventRecord(c->ev[0], sp));
      // its triangle's inverse, beside the next step's interchanges
      if (n > jn + jbn) inv_tri(jn, jbn, sp, pp, flp);
      LA_TRY(hipEventRecord(c->ev[4], sp));
    }
    if (jbn > 0) {
      mark(pa, fla, LA_PHASE_TRSM, 0);
      K<T>::copy(jbn, jb, X, ldx, Tp(jn, j), ldt, s, npx);
      endm(pa, fla);
    }
    // ---- the rest, beside the next panel: interchanges left of this panel
    //      and right of the next one, their TRSM and update
    if (!(pre && jbn > 0 && mrows > 0)) rest_swap_trsm(npx);
    if (nr > 0) {
      if (mrows > 0) {
        mark(pa, fla, LA_PHASE_GEMM, 2.0 * nr * (double)mrows * jb);
        K<T>::update(nr, mrows, jb, X + jbn, ldx, Tp(j, jn), ldt, Tp(jn + jbn, jn), ldt, npx, s);
        endm(pa, fla);
      }
    }
  }
  if (rc) {
    // drain the panel stream before giving the workspace back
    hipStreamSynchronize(sp);
    hipStreamSynchronize(s);
    ws_release(c, s_user);
    return rc;
  }
  // the factors go back to A only when no bounded wait gave up: otherwise A
  // is left as the caller passed it, INFO = LA_ERR_SYNC, and the call can be
  // redone
  mark(pa, fla, LA_PHASE_TRANSPOSE, 0);
  K<T>::transpose(n, m, Tm, ldt, A, lda, s, err);
  endm(pa, fla);
  launch_fold_err(err, info, c->status, s);
  LA_TRY(hipEventRecord(c->ev[3], s));
  LA_TRY(hipStreamWaitEvent(s_user, c->ev[3], 0));
  rc = ws_release(c, s_user);
  phase_collect(pa, fla);
  phase_collect(pp, flp);
  return rc;
}

// --------------------------------------------------------------- GETRS --

// Workspace of the chain form of DGETRS (getrs_chain).
template <class T>
size_t getrs_chain_ws_bytes(int n, int nrhs) {
  Carve cv;
  cv.take<int>((size_t)n);
  cv.take<T>((size_t)n * nrhs);
  cv.take<T>((size_t)n * nrhs);
  cv.take<unsigned>(tri_chain_ctl_words(nrhs));
  cv.take<int>(1);
  cv.take<int>(perm_pairs_ints(0, n));
  return cv.off;
}

// Workspace of the blocked form (a redone call).
template <class T>
size_t getrs_ws_bytes(int n, int nrhs, int nbs) {
  Carve cv;
  cv.take<int>((size_t)n);
  cv.take<T>((size_t)n * nrhs);
  cv.take<T>((size_t)nbs * nrhs);
  cv.take<T>((size_t)nbs * (size_t)((n + nbs - 1) / nbs) * nbs);
  cv.take<T>((size_t)nbs * (size_t)((n + nbs - 1) / nbs) * nbs);
  cv.take<int>(perm_pairs_ints(0, n));
  return std::max(cv.off, getrs_chain_ws_bytes<T>(n, nrhs));
}

// The chain form of DGETRS (SRC/dgetrs.f:181-218) on the device workspace
// ws (getrs_chain_ws_bytes(n, nrhs) bytes, 256-byte aligned): the row
// permutation, then one tri_chain launch per triangular sweep (getrs.hip).
// When a wait gives up, B is left as passed and the failure is folded into
// *waitflag and *sticky (device ints; either may be null).  fail: start with
// the wait-error word set (the test hook).  Enqueued on s; allocates nothing.
template <class T>
void getrs_chain(char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb, char* ws,
                 bool fail, int* waitflag, int* sticky, hipStream_t s) {
  const bool notran = (trans == 'N' || trans == 'n');
  Carve cv;
  int* perm = reinterpret_cast<int*>(ws + cv.take<int>((size_t)n));
  T* X = reinterpret_cast<T*>(ws + cv.take<T>((size_t)n * nrhs));
  T* Y = reinterpret_cast<T*>(ws + cv.take<T>((size_t)n * nrhs));
  unsigned* ctl = reinterpret_cast<unsigned*>(ws + cv.take<unsigned>(tri_chain_ctl_words(nrhs)));
  int* cerr = reinterpret_cast<int*>(ws + cv.take<int>(1));
  int* gpairs = reinterpret_cast<int*>(ws + cv.take<int>(perm_pairs_ints(0, n)));
  const long ldx = n;
  launch_fill_i32(cerr, 1, fail ? 1 : 0, s);
  if (notran) {
    launch_swaps_to_perm(n, 0, n, ipiv, 1, perm, s, gpairs);
    K<T>::gather(n, nrhs, perm, B, ldb, X, ldx, s);                         // X = P*B (dgetrs.f:187)
    launch_tri_chain<T>(0, n, nrhs, A, lda, X, ldx, Y, ldx, ctl, cerr, s);  // L Y = X (dgetrs.f:191)
    launch_tri_chain<T>(1, n, nrhs, A, lda, Y, ldx, X, ldx, ctl, cerr, s);  // U X = Y (dgetrs.f:197)
    K<T>::copy(n, nrhs, X, ldx, B, ldb, s, 0, cerr);
  } else {
    K<T>::copy(n, nrhs, B, ldb, X, ldx, s);
    launch_tri_chain<T>(2, n, nrhs, A, lda, X, ldx, Y, ldx, ctl, cerr, s);  // U^T Y = B (dgetrs.f:205)
    launch_tri_chain<T>(3, n, nrhs, A, lda, Y, ldx, X, ldx, ctl, cerr, s);  // L^T X = Y (dgetrs.f:211)
    launch_swaps_to_perm(n, 0, n, ipiv, -1, perm, s, gpairs);                // dgetrs.f:217
    K<T>::gather(n, nrhs, perm, X, ldx, B, ldb, s, cerr);
  }
  launch_fold_err(cerr, waitflag, sticky, s);
}

// Solve op(A) X = B with the LU of getrf (SRC/dgetrs.f:181-218).  Device
// pointers, asynchronous on s.  Two forms of the triangular sweeps:
//   * getrs_chain, the default: one tri_chain launch per sweep (getrs.hip),
//     workgroups in ticket order with bounded waits for the blocks they
//     depend on.  When a wait gives up, B is left as passed, bit 0 of the
//     device's sticky status is set and, when waitflag is given, so is
//     *waitflag (a device int), so that a synchronous caller can redo the
//     solve;
//   * the blocked form without waits (t_safe, a redone call): the diagonal
//     blocks of L and U inverted once, every block step two MFMA GEMMs.
template <class T>
int getrs_device_ex(char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb,
                    hipStream_t s, int* waitflag) {
  DevCtx* c = get_ctx();
  if (!c) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(c->mu);
  if (n == 0 || nrhs == 0) return 0;
  const bool notran = (trans == 'N' || trans == 'n');
  const int nbs = 256;
  const int nblk = (n + nbs - 1) / nbs;
  char* ws = nullptr;
  int rc = ws_acquire(c, t_safe ? getrs_ws_bytes<T>(n, nrhs, nbs) : getrs_chain_ws_bytes<T>(n, nrhs), s, &ws);
  if (rc) return rc;

  PhaseTimer pt;
  pt.on = g_phase_on.load() != 0;
  pt.level = g_phase_on.load();
  pt.s = s;
  std::vector<double> fl;
  if (pt.mark(LA_PHASE_GETRS)) fl.push_back(2.0 * (double)n * n * nrhs);

  if (!t_safe) {
    int fail = g_fail_waits.load();
    while (fail > 0 && !g_fail_waits.compare_exchange_weak(fail, fail - 1)) {
    }
    getrs_chain<T>(trans, n, nrhs, A, lda, ipiv, B, ldb, ws, fail > 0, waitflag, c->status, s);
    if (pt.mark(-1)) fl.push_back(0);
    rc = ws_release(c, s);
    phase_collect(pt, fl);
    return rc;
  }
  Carve cv;
  int* perm = reinterpret_cast<int*>(ws + cv.take<int>((size_t)n));
  T* X = reinterpret_cast<T*>(ws + cv.take<T>((size_t)n * nrhs));
  T* Y = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nbs * nrhs));
  T* WL = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nbs * nblk * nbs));
  T* WU = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nbs * nblk * nbs));
  int* gpairs = reinterpret_cast<int*>(ws + cv.take<int>(perm_pairs_ints(0, n)));
  const long ldx = n, ldy = nbs, ldw = nbs;
  auto Ap = [&](long r, long col) { return A + r + col * lda; };
  K<T>::trtri_blocks(false, true, n, nbs, A, lda, WL, ldw, s);   // inv(L_dd), unit lower
  K<T>::trtri_blocks(true, false, n, nbs, A, lda, WU, ldw, s);   // inv(U_dd), non-unit upper
  if (notran) {
    launch_swaps_to_perm(n, 0, n, ipiv, 1, perm, s, gpairs);
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
    launch_swaps_to_perm(n, 0, n, ipiv, -1, perm, s, gpairs);              // dgetrs.f:217
    K<T>::gather(n, nrhs, perm, X, ldx, B, ldb, s);
  }
  if (pt.mark(-1)) fl.push_back(0);
  rc = ws_release(c, s);
  phase_collect(pt, fl);
  return rc;
}

template <class T>
int getrs_device(char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb,
                 hipStream_t s) {
  return getrs_device_ex<T>(trans, n, nrhs, A, lda, ipiv, B, ldb, s, nullptr);
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
  // The whole sequence is composed into one permutation (blocks of 256
  // interchanges composed in parallel), which moves each column's rows in
  // place through LDS in one pass; with more rows than LDS holds, the rows
  // are gathered into scratch and copied back.
  const bool inplace = nrows <= permute_rows_max((int)sizeof(T));
  char* ws = nullptr;
  Carve cv0;
  cv0.take<int>((size_t)nrows);
  cv0.take<int>(perm_pairs_ints(k1 - 1, k2));
  if (!inplace) cv0.take<T>((size_t)nrows * ncols);
  int rc = ws_acquire(c, cv0.off, s, &ws);
  if (rc) return rc;
  Carve cv;
  int* perm = reinterpret_cast<int*>(ws + cv.take<int>((size_t)nrows));
  int* gpairs = reinterpret_cast<int*>(ws + cv.take<int>(perm_pairs_ints(k1 - 1, k2)));
  // k1, k2 are 1-based and inclusive in the reference; perm kernel takes
  // 0-based [k1-1, k2).
  launch_swaps_to_perm(nrows, k1 - 1, k2, ipiv, incx, perm, s, gpairs);
  if (inplace) {
    K<T>::permute_rows(nrows, ncols, perm, A, lda, s);
  } else {
    T* X = reinterpret_cast<T*>(ws + cv.take<T>((size_t)nrows * ncols));
    K<T>::gather(nrows, ncols, perm, A, lda, X, nrows, s);
    K<T>::copy(nrows, ncols, X, nrows, A, lda, s);
  }
  return ws_release(c, s);
}

// FP32-stored LU whose trailing updates take FP16 (kind 1) or BF16 (kind 2)
// operands with FP32 accumulation; kind 0 is plain SGETRF.
int getrf_device_lp(int m, int n, float* A, long lda, int* ipiv, int* info, int kind, hipStream_t s) {
  t_upd_kind = kind;
  const int rc = getrf_device<float>(m, n, A, lda, ipiv, info, s);
  t_upd_kind = 0;
  return rc;
}

// Explicit instantiations used by the C ABI and the mixed-precision solver.
template int getrf_device<double>(int, int, double*, long, int*, int*, hipStream_t);
template int getrf_device<float>(int, int, float*, long, int*, int*, hipStream_t);
template int getrs_device<double>(char, int, int, const double*, long, const int*, double*, long, hipStream_t);
template int getrs_device<float>(char, int, int, const float*, long, const int*, float*, long, hipStream_t);
template int getrs_device_ex<double>(char, int, int, const double*, long, const int*, double*, long, hipStream_t,
                                     int*);
template int getrs_device_ex<float>(char, int, int, const float*, long, const int*, float*, long, hipStream_t,
                                    int*);
template size_t getrs_chain_ws_bytes<double>(int, int);
template void getrs_chain<double>(char, int, int, const double*, long, const int*, double*, long, char*, bool, int*,
                                  int*, hipStream_t);
template int laswp_device<double>(int, double*, long, int, int, int, const int*, int, hipStream_t);
template int laswp_device<float>(int, float*, long, int, int, int, const int*, int, hipStream_t);

int last_error() { return t_last_error; }

// Autotuner of the plain factorization: times DGETRF of an n x n matrix
// (uniform(-1,1) entries from a device generator) for each candidate
// (NB, leaf) -- one untimed call, then the best of two timed ones, HIP
// events on the library's stream -- and keeps the fastest for n's bucket
// (the next power of two); with path, the table is written there as JSON
// (and read back by lapack_amd_load_tuning or LAPACK_AMD_TUNE_FILE).
int tune_getrf(int n, const char* path, int* nb_out, int* leaf_out, double* ms_out) {
  if (n <= 0) return -1;
  DevCtx* c = get_ctx();
  if (!c) return LA_ERR_NODEVICE;
  const size_t bytes = sizeof(double) * (size_t)n * n;
  double *A0 = nullptr, *A = nullptr;
  int *ipiv = nullptr, *info = nullptr;
  if (hipMalloc(&A0, bytes) != hipSuccess || hipMalloc(&A, bytes) != hipSuccess ||
      hipMalloc(&ipiv, sizeof(int) * (size_t)n + 256) != hipSuccess) {
    (void)hipGetLastError();
    if (A0) hipFree(A0);
    if (A) hipFree(A);
    return LA_ERR_NOMEM;
  }
  info = ipiv + n;
  hipStream_t s = c->stream;
  launch_fill_random(A0, (size_t)n * n, 0x9e3779b97f4a7c15ull ^ (unsigned long long)n, s);
  static const int cand[][2] = {{256, 64}, {192, 64}, {128, 64}, {256, 32}, {192, 32}, {128, 32}};
  hipEvent_t e0, e1;
  int rc = 0;
  TuneEntry best;
  best.ms = 1e300;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = LA_ERR_HIP;
  for (const auto& cd : cand) {
    if (rc) break;
    t_force_nb = cd[0];
    t_force_leaf = cd[1];
    double tmin = 1e300;
    for (int rep = 0; rep < 3 && !rc; ++rep) {
      if (hipMemcpyAsync(A, A0, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) { rc = LA_ERR_HIP; break; }
      hipEventRecord(e0, s);
      rc = getrf_device<double>(n, n, A, n, ipiv, info, s);
      hipEventRecord(e1, s);
      if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = LA_ERR_HIP;
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0) tmin = std::min(tmin, (double)ms);
    }
    int h_info = 0;
    if (!rc) hipMemcpy(&h_info, info, sizeof(int), hipMemcpyDeviceToHost);
    if (!rc && h_info == 0 && tmin < best.ms) {
      best.nb = cd[0];
      best.leaf = cd[1];
      best.ms = tmin;
    }
  }
  t_force_nb = 0;
  t_force_leaf = 0;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(A0);
  hipFree(A);
  hipFree(ipiv);
  if (rc) return rc;
  if (best.nb <= 0) return LA_ERR_SYNC;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_file_read = true;
    if (path && *path) (void)load_tuning_locked(path);  // keep the other buckets
    g_tune[tune_bucket(n)] = best;
    if (path && *path && save_tuning_locked(path) != 0) return LA_ERR_HIP;
  }
  if (nb_out) *nb_out = best.nb;
  if (leaf_out) *leaf_out = best.leaf;
  if (ms_out) *ms_out = best.ms;
  return 0;
}

// Runs f() with the forms without waits selected (a redone call).
int with_safe_forms(const std::function<int()>& f) {
  const int old = t_safe;
  t_safe = 1;
  const int rc = f();
  t_safe = old;
  return rc;
}
void fail_waits(int count) { g_fail_waits = count > 0 ? count : 0; }
// Bit 0: a bounded wait of a kernel launched on the current device gave up
// since the last clear.  Waits for the device to be idle first.
int device_status(int clear) {
  DevCtx* c = get_ctx();
  if (!c) return LA_ERR_NODEVICE;
  LA_TRY(hipDeviceSynchronize());
  int v = 0;
  LA_TRY(hipMemcpy(&v, c->status, sizeof(int), hipMemcpyDeviceToHost));
  if (clear && v) LA_TRY(hipMemset(c->status, 0, sizeof(int)));
  return v;
}
void set_tuning(int nb, int leaf) { g_nb = nb; g_leaf = leaf; }
DevCtx* ctx_for_current() { return get_ctx(); }
hipStream_t ctx_stream(DevCtx* c) { return c->stream; }
int io_buffer(DevCtx* c, size_t bytes, char** out) { return io_acquire(c, bytes, out); }
std::mutex& io_mutex(DevCtx* c) { return c->io_mu; }
int mp_buffer(DevCtx* c, size_t bytes, char** out) { return mp_acquire(c, bytes, out); }
std::mutex& mp_mutex(DevCtx* c) { return c->mp_mu; }
void phase_enable(int on) { g_phase_on = on; }
void phase_read(double* ms, long* calls, double* flops) {
  std::lock_guard<std::mutex> g(g_phase_mu);
  phase_resolve_locked();
  for (int i = 0; i < LA_PHASE_COUNT; ++i) {
    if (ms) ms[i] = g_phase_ms[i];
    if (calls) calls[i] = g_phase_calls[i];
    if (flops) flops[i] = g_phase_flops[i];
  }
}
void phase_reset() {
  std::lock_guard<std::mutex> g(g_phase_mu);
  phase_resolve_locked();
  for (int i = 0; i < LA_PHASE_COUNT; ++i) { g_phase_ms[i] = 0; g_phase_calls[i] = 0; g_phase_flops[i] = 0; }
}

}  // namespace la
