// noalloc.hip -- DGETRS on a static device scratch array, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops
// that forbid hipMalloc).
//
// lapack_amd_dgetrs_noalloc_async runs the same kernels as
// lapack_amd_dgetrs_async (the chain form of DGETRS, SRC/dgetrs.f:181-218),
// but its workspace -- the row permutation, the two n x nrhs buffers the
// out-of-place sweeps alternate between, the sweeps' ticket words, the
// wait-error word and the composed interchange pairs -- lives in `scratch`,
// a file-scope device array of LA_NOALLOC_SCRATCH_BYTES per GPU (held until
// the process ends).  Host code never passes the array's name to a kernel:
// the first call on a device takes the array's device address with
// hipGetSymbolAddress (with that device current) and keeps it per device.
// Every call on one GPU shares that GPU's array, so calls that overlap on
// one GPU (on different streams, or from different threads without
// ordering) race; calls on one stream are ordered.  The scratch needs no
// clearing between calls: every word a call reads it has written itself
// first (before each sweep, fill kernels zero its ticket words and set its
// output to the sentinel the sweep polls for).  No hipMalloc, hipFree or
// allocating library routine is called.
#include "la_common.h"

#include <atomic>
#include <mutex>

namespace la {
int getrs_arg_check(char trans, int n, int nrhs, int lda, int ldb);
template <class T>
size_t getrs_chain_ws_bytes(int n, int nrhs);
template <class T>
void getrs_chain(char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb, char* ws,
                 bool fail, int* waitflag, int* sticky, hipStream_t s);
}  // namespace la

namespace {

constexpr size_t LA_NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;  // 16 MiB per GPU
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) char scratch[LA_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::atomic<char*> g_scratch_ptr[MAX_DEVICES];

// The device address of `scratch` on the current device (taken once per
// device), or null with *rc set.
char* scratch_on_current_device(int* rc) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    *rc = LA_ERR_NODEVICE;
    return nullptr;
  }
  if (dev < 0 || dev >= MAX_DEVICES) {
    *rc = LA_ERR_LIMIT;
    return nullptr;
  }
  char* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (p) return p;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
  if (!p) {
    void* addr = nullptr;
    if (hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)) != hipSuccess || !addr) {
      (void)hipGetLastError();
      *rc = LA_ERR_HIP;
      return nullptr;
    }
    p = static_cast<char*>(addr);
    g_scratch_ptr[dev].store(p, std::memory_order_release);
  }
  return p;
}

}  // namespace

extern "C" {

size_t lapack_amd_dgetrs_noalloc_bytes(int n, int nrhs) {
  return (n < 0 || nrhs < 0) ? 0 : la::getrs_chain_ws_bytes<double>(n, nrhs);
}

size_t lapack_amd_noalloc_scratch_bytes(void) { return LA_NOALLOC_SCRATCH_BYTES; }

int lapack_amd_dgetrs_noalloc_async(char trans, int n, int nrhs, const double* d_a, int lda, const int* d_ipiv,
                                    double* d_b, int ldb, int* d_info, void* stream) {
  int rc = la::getrs_arg_check(trans, n, nrhs, lda, ldb);
  if (rc) return rc;
  if (!d_info) return -9;
  if (la::getrs_chain_ws_bytes<double>(n, nrhs) > LA_NOALLOC_SCRATCH_BYTES) return LA_ERR_LIMIT;
  hipStream_t s = (hipStream_t)stream;
  char* ws = scratch_on_current_device(&rc);
  if (!ws) return rc;
  la::launch_fill_i32(d_info, 1, 0, s);
  if (n > 0 && nrhs > 0)
    la::getrs_chain<double>(trans, n, nrhs, d_a, lda, d_ipiv, d_b, ldb, ws, false, d_info, nullptr, s);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : LA_ERR_HIP;
}

}  // extern "C"
