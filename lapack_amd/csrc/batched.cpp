// batched.cpp -- plans, autotuner and C entry points of the batched LU.
//
// A plan fixes a batch's shape: the order n of every matrix (host array at
// creation), the device copy of it, the order in which workgroups take the
// matrices (decreasing n) and the kernel configuration (panel width NB and
// workgroup size) chosen by the autotuner.  Creating a plan allocates and
// may run kernels; the launches that use it allocate nothing, synchronize
// nothing and make only kernel launches, so they can be captured in a
// hipGraph and replayed.
//
// Autotuner: for each configuration that fits the device's LDS at the
// batch's largest n, the plan's own size distribution is factored on
// synthetic uniform(-1,1) matrices (a deterministic sample of up to 2048 of
// the batch's sizes) and timed with HIP events; the fastest is kept, and
// remembered per (precision, nmax, mean n) for later plans.
#include "la_common.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

#include "../../include/lapack_amd.h"

namespace la {
namespace {

int hipc(hipError_t e) {
  if (e == hipSuccess) return 0;
  std::fprintf(stderr, "lapack_amd: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
  return LA_ERR_HIP;
}
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, int> g_tuned;  // (elem, nmax, mean n / 16) -> config

}  // namespace
}  // namespace la

struct lapack_amd_batch_plan {
  int batch = 0;
  int nmax = 0;
  int elem = 8;     // 8: FP64, 4: FP32
  int cfg = -1;     // kernel configuration index
  int grid_f = 1;   // persistent grid of the getrf launch
  int grid_s = 1;   // persistent grid of the getrs launch
  int grid_t = 1;   // persistent grid of the adjoint's transposed solve (FP64)
  int dev = -1;
  int* d_n = nullptr;
  int* d_order = nullptr;
  unsigned* d_counter = nullptr;  // [0] getrf, [1] getrs, [2] adjoint
  std::vector<int> n;
  double tuned_ms = 0.0;
};

namespace la {
namespace {

template <class T>
int tune_plan(lapack_amd_batch_plan* p, hipStream_t s) {
  // Sample of the batch's sizes: every k-th matrix of the decreasing order.
  std::vector<int> sorted = p->n;
  std::sort(sorted.begin(), sorted.end(), std::greater<int>());
  const int cap = 2048;
  const int step = std::max(1, (int)((sorted.size() + cap - 1) / cap));
  std::vector<int> ns;
  for (size_t i = 0; i < sorted.size(); i += step) ns.push_back(sorted[i]);
  const int nb = (int)ns.size();
  std::vector<size_t> off(nb + 1, 0);
  for (int i = 0; i < nb; ++i) off[i + 1] = off[i] + (size_t)ns[i] * ns[i];
  const size_t total = std::max<size_t>(off[nb], 1);
  T *pristine = nullptr, *work = nullptr;
  T** d_ptr = nullptr;
  int *d_ns = nullptr, *d_lda = nullptr, *d_ipiv = nullptr, *d_info = nullptr, *d_ord = nullptr;
  int** d_ipp = nullptr;
  unsigned* d_ctr = nullptr;
  int rc = 0;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    for (void* q : {(void*)pristine, (void*)work, (void*)d_ptr, (void*)d_ns, (void*)d_lda, (void*)d_ipiv,
                    (void*)d_info, (void*)d_ord, (void*)d_ipp, (void*)d_ctr})
      if (q) (void)hipFree(q);
  };
  if (hipMalloc(&pristine, sizeof(T) * total) != hipSuccess || hipMalloc(&work, sizeof(T) * total) != hipSuccess ||
      hipMalloc(&d_ptr, sizeof(T*) * nb) != hipSuccess || hipMalloc(&d_ns, sizeof(int) * nb) != hipSuccess ||
      hipMalloc(&d_lda, sizeof(int) * nb) != hipSuccess ||
      hipMalloc(&d_ipiv, sizeof(int) * std::max<size_t>(1, (size_t)nb * p->nmax)) != hipSuccess ||
      hipMalloc(&d_info, sizeof(int) * nb) != hipSuccess || hipMalloc(&d_ord, sizeof(int) * nb) != hipSuccess ||
      hipMalloc(&d_ipp, sizeof(int*) * nb) != hipSuccess || hipMalloc(&d_ctr, 256) != hipSuccess) {
    (void)hipGetLastError();
    cleanup();
    return LA_ERR_NOMEM;
  }
  std::vector<T*> hp(nb);
  std::vector<int*> hip_(nb);
  std::vector<int> ord(nb);
  for (int i = 0; i < nb; ++i) {
    hp[i] = work + off[i];
    hip_[i] = d_ipiv + (size_t)i * p->nmax;
    ord[i] = i;
  }
  if ((rc = hipc(hipMemcpyAsync(d_ptr, hp.data(), sizeof(T*) * nb, hipMemcpyHostToDevice, s))) ||
      (rc = hipc(hipMemcpyAsync(d_ns, ns.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s))) ||
      (rc = hipc(hipMemcpyAsync(d_lda, ns.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s))) ||
      (rc = hipc(hipMemcpyAsync(d_ord, ord.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s))) ||
      (rc = hipc(hipMemcpyAsync(d_ipp, hip_.data(), sizeof(int*) * nb, hipMemcpyHostToDevice, s)))) {
    cleanup();
    return rc;
  }
  if (sizeof(T) == 8)
    launch_fill_uniform_d(reinterpret_cast<double*>(pristine), total, 0x5EEDull, s);
  else
    launch_fill_uniform_s(reinterpret_cast<float*>(pristine), total, 0x5EEDull, s);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int best = -1;
  float best_ms = 1e30f;
  for (int cfg = 0; cfg < batched_getrf_configs(); ++cfg) {
    if (batched_getrf_smem(cfg, p->nmax, (int)sizeof(T)) > (size_t)lds_limit()) continue;
    float ms_min = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      (void)hipMemcpyAsync(work, pristine, sizeof(T) * total, hipMemcpyDeviceToDevice, s);
      (void)hipEventRecord(e0, s);
      const int grid = batched_getrf_grid<T>(cfg, nb, p->nmax);
      if (grid < 0) break;
      int r = launch_getrf_batched<T>(cfg, grid, nb, p->nmax, d_ns, d_ptr, d_lda, d_ipp, d_info, d_ord, d_ctr, s);
      (void)hipEventRecord(e1, s);
      if (r) break;
      if (hipEventSynchronize(e1) != hipSuccess) { (void)hipGetLastError(); break; }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms_min = std::min(ms_min, ms);
    }
    if (ms_min < best_ms) { best_ms = ms_min; best = cfg; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  cleanup();
  if (best < 0) return LA_ERR_LIMIT;
  p->cfg = best;
  p->tuned_ms = best_ms;
  return 0;
}

int default_cfg(int nmax, int elem) {
  // Largest panel that fits, 256 threads (autotuning usually confirms NB 16 or 32).
  for (int cfg : {3, 1, 0, 5})
    if (batched_getrf_smem(cfg, nmax, elem) <= (size_t)lds_limit()) return cfg;
  return -1;
}

}  // namespace
}  // namespace la

using namespace la;

// Persistent grids for the plan's configuration (queries the runtime once).
static int rc_grid(lapack_amd_batch_plan* p) {
  const int nm = std::max(p->nmax, 1), b = std::max(p->batch, 1);
  const int gf = p->elem == 8 ? batched_getrf_grid<double>(p->cfg, b, nm) : batched_getrf_grid<float>(p->cfg, b, nm);
  const int gs = p->elem == 8 ? batched_getrs_grid<double>(b, nm) : batched_getrs_grid<float>(b, nm);
  if (gf < 0) return gf;
  if (gs < 0) return gs;
  p->grid_f = gf;
  p->grid_s = gs;
  // the adjoint's grid; a plan whose orders exceed its LDS keeps 0 (the
  // adjoint then reports LA_ERR_LIMIT)
  p->grid_t = p->elem == 8 ? std::max(0, batched_getrs_t_grid<double>(b, nm)) : 0;
  return 0;
}

extern "C" {

int lapack_amd_batch_plan_create(lapack_amd_batch_plan** out, int batch, const int* n, char precision,
                                 int autotune) {
  if (!out) return -1;
  *out = nullptr;
  if (batch < 0) return -2;
  if (!n && batch > 0) return -3;
  const bool dbl = (precision == 'd' || precision == 'D');
  if (!dbl && precision != 's' && precision != 'S') return -4;
  auto* p = new lapack_amd_batch_plan();
  p->batch = batch;
  p->elem = dbl ? 8 : 4;
  p->n.assign(n, n + batch);
  for (int v : p->n)
    if (v < 0) { delete p; return -3; }
  p->nmax = batch ? *std::max_element(p->n.begin(), p->n.end()) : 0;
  if (hipGetDevice(&p->dev) != hipSuccess) { (void)hipGetLastError(); delete p; return LA_ERR_NODEVICE; }
  std::vector<int> order(batch);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return p->n[a] > p->n[b]; });
  const size_t nbytes = sizeof(int) * std::max(batch, 1);
  if (hipMalloc(&p->d_n, nbytes) != hipSuccess || hipMalloc(&p->d_order, nbytes) != hipSuccess ||
      hipMalloc(&p->d_counter, 256) != hipSuccess) {
    (void)hipGetLastError();
    lapack_amd_batch_plan_destroy(p);
    return LA_ERR_NOMEM;
  }
  if (batch > 0 && (hipMemcpy(p->d_n, p->n.data(), nbytes, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(p->d_order, order.data(), nbytes, hipMemcpyHostToDevice) != hipSuccess)) {
    (void)hipGetLastError();
    lapack_amd_batch_plan_destroy(p);
    return LA_ERR_HIP;
  }
  p->cfg = default_cfg(std::max(p->nmax, 1), p->elem);
  if (p->cfg < 0) { lapack_amd_batch_plan_destroy(p); return LA_ERR_LIMIT; }
  if (autotune && batch > 0) {
    long long sum = 0;
    for (int v : p->n) sum += v;
    const auto key = std::make_tuple(p->elem, p->nmax, (int)(sum / batch / 16));
    int cached = -1;
    {
      std::lock_guard<std::mutex> g(g_tune_mu);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) cached = it->second;
    }
    if (cached >= 0) {
      p->cfg = cached;
    } else {
      hipStream_t s = nullptr;
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        lapack_amd_batch_plan_destroy(p);
        return LA_ERR_HIP;
      }
      int rc = dbl ? tune_plan<double>(p, s) : tune_plan<float>(p, s);
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
      if (rc) { lapack_amd_batch_plan_destroy(p); return rc; }
      std::lock_guard<std::mutex> g(g_tune_mu);
      g_tuned[key] = p->cfg;
    }
  }
  const int rg = rc_grid(p);
  if (rg < 0) { lapack_amd_batch_plan_destroy(p); return rg; }
  *out = p;
  return 0;
}

void lapack_amd_batch_plan_destroy(lapack_amd_batch_plan* p) {
  if (!p) return;
  (void)hipDeviceSynchronize();
  (void)hipFree(p->d_n);
  (void)hipFree(p->d_order);
  (void)hipFree(p->d_counter);
  delete p;
}

int lapack_amd_batch_plan_config(const lapack_amd_batch_plan* p, int* nb, int* nthreads, double* tuned_ms) {
  if (!p) return -1;
  int b = 0, t = 0;
  batched_getrf_config(p->cfg, &b, &t);
  if (nb) *nb = b;
  if (nthreads) *nthreads = t;
  if (tuned_ms) *tuned_ms = p->tuned_ms;
  return 0;
}

int lapack_amd_batch_plan_set_config(lapack_amd_batch_plan* p, int nb, int nthreads) {
  if (!p) return -1;
  for (int c = 0; c < batched_getrf_configs(); ++c) {
    int b, t;
    batched_getrf_config(c, &b, &t);
    if (b == nb && t == nthreads) {
      if (batched_getrf_smem(c, std::max(p->nmax, 1), p->elem) > (size_t)lds_limit()) return LA_ERR_LIMIT;
      const int old = p->cfg;
      p->cfg = c;
      const int r = rc_grid(p);
      if (r < 0) p->cfg = old;
      return r;
    }
  }
  return -2;
}

int lapack_amd_dgetrf_batched_async(const lapack_amd_batch_plan* p, double* const* d_a, const int* d_lda,
                                    int* const* d_ipiv, int* d_info, void* stream) {
  if (!p) return -1;
  if (p->elem != 8) return -1;
  if (p->batch == 0) return 0;
  int rc = launch_getrf_batched<double>(p->cfg, p->grid_f, p->batch, p->nmax, p->d_n, d_a, d_lda, d_ipiv, d_info,
                                        p->d_order,
                                        p->d_counter, (hipStream_t)stream);
  return rc ? rc : hipc(hipGetLastError());
}
int lapack_amd_sgetrf_batched_async(const lapack_amd_batch_plan* p, float* const* d_a, const int* d_lda,
                                    int* const* d_ipiv, int* d_info, void* stream) {
  if (!p) return -1;
  if (p->elem != 4) return -1;
  if (p->batch == 0) return 0;
  int rc = launch_getrf_batched<float>(p->cfg, p->grid_f, p->batch, p->nmax, p->d_n, d_a, d_lda, d_ipiv, d_info,
                                       p->d_order,
                                       p->d_counter, (hipStream_t)stream);
  return rc ? rc : hipc(hipGetLastError());
}
int lapack_amd_dgetrs_batched_async(const lapack_amd_batch_plan* p, const int* d_nrhs, double* const* d_a,
                                    const int* d_lda, int* const* d_ipiv, double* const* d_b, const int* d_ldb,
                                    void* stream) {
  if (!p) return -1;
  if (p->elem != 8) return -1;
  if (p->batch == 0) return 0;
  int rc = launch_getrs_batched<double>(p->grid_s, p->batch, p->nmax, p->d_n, d_nrhs, d_a, d_lda, d_ipiv, d_b, d_ldb,
                                        p->d_order, p->d_counter + 1, (hipStream_t)stream);
  return rc ? rc : hipc(hipGetLastError());
}
int lapack_amd_sgetrs_batched_async(const lapack_amd_batch_plan* p, const int* d_nrhs, float* const* d_a,
                                    const int* d_lda, int* const* d_ipiv, float* const* d_b, const int* d_ldb,
                                    void* stream) {
  if (!p) return -1;
  if (p->elem != 4) return -1;
  if (p->batch == 0) return 0;
  int rc = launch_getrs_batched<float>(p->grid_s, p->batch, p->nmax, p->d_n, d_nrhs, d_a, d_lda, d_ipiv, d_b, d_ldb,
                                       p->d_order, p->d_counter + 1, (hipStream_t)stream);
  return rc ? rc : hipc(hipGetLastError());
}


/* Batched adjoint of the solve (see the header). */
int lapack_amd_dgesv_adjoint_batched_async(const lapack_amd_batch_plan* p, const int* d_nrhs, double* const* d_lu,
                                           const int* d_lda, int* const* d_ipiv, double* const* d_x, const int* d_ldx,
                                           double* const* d_xbar, const int* d_ldxbar, double* const* d_abar,
                                           const int* d_ldabar, int accumulate, void* stream) {
  if (!p) return -1;
  if (p->elem != 8) return -1;
  if (p->batch == 0) return 0;
  if (p->grid_t <= 0) return LA_ERR_LIMIT;
  int rc = launch_adjoint_batched(p->grid_t, p->batch, p->nmax, p->d_n, d_nrhs, d_lu, d_lda, d_ipiv, d_x, d_ldx,
                                  d_xbar, d_ldxbar, d_abar, d_ldabar, accumulate, p->d_order, p->d_counter + 2,
                                  (hipStream_t)stream);
  return rc ? rc : hipc(hipGetLastError());
}

}  // extern "C"
