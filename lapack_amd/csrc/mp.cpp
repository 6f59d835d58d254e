// mp.cpp -- DSGESV on one gfx950 GPU: FP32 LU on the FP32 matrix cores, FP64
// iterative refinement, FP64 LU fallback (SRC/dsgesv.f:259-427).
//
// The sequence of operations and every decision follow the reference:
//   CTE = DLANGE('I',A) * DLAMCH('Epsilon') * SQRT(N) * BWDMAX  (dsgesv.f:286-288)
//   SX = DLAG2S(B), SA = DLAG2S(A)     overflow        -> ITER = -2 (dsgesv.f:292-305)
//   SGETRF(SA)                         INFO /= 0       -> ITER = -3 (dsgesv.f:309-314)
//   X = SLAG2D(SGETRS(SA, SX)); R = B - A*X            (dsgesv.f:318-332)
//   stop when max|R(:,i)| <= max|X(:,i)|*CTE for every column i (IDAMAX norms,
//   dsgesv.f:337-342), else up to ITERMAX = 30 corrections
//   X += SLAG2D(SGETRS(SA, DLAG2S(R)))  (dsgesv.f:352-390); never -> ITER = -31
//   fallback: DGETRF(A), X = B, DGETRS  (dsgesv.f:413-427).
// The residual, the update and the norms stay in FP64 on the GPU; only the
// per-column norms (2*NRHS numbers) and three flags travel to the host, where
// the comparisons are made with the reference's exact expressions.
#include "la_common.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace la {
struct DevCtx;
template <class T> int getrf_device(int m, int n, T* A, long lda, int* ipiv, int* info, hipStream_t s);
template <class T>
int getrs_device(char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb, hipStream_t s);
DevCtx* ctx_for_current();
int mp_buffer(DevCtx* c, size_t bytes, char** out);
std::mutex& mp_mutex(DevCtx* c);

namespace {

size_t al256(size_t x) { return (x + 255) / 256 * 256; }

int hipc(hipError_t e) {
  if (e == hipSuccess) return 0;
  std::fprintf(stderr, "lapack_amd: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
  return LA_ERR_HIP;
}
#define MP_CK(x)          \
  do {                    \
    int _r = hipc(x);     \
    if (_r) return _r;    \
  } while (0)

double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

}  // namespace

// All arrays are device memory.  Returns INFO (>= 0) or a negative lapack_amd
// error code; *iter receives ITER.  Synchronous: the host takes each
// decision of the reference from 2*nrhs norms copied back per step.
int dsgesv_device(int n, int nrhs, double* A, long lda, int* ipiv, const double* B, long ldb, double* X, long ldx,
                  int* iter, hipStream_t s) {
  *iter = 0;
  if (n == 0) return 0;
  DevCtx* c = ctx_for_current();
  if (!c) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(mp_mutex(c));
  const int itermax = 30;
  const int nr = nrhs > 0 ? nrhs : 1;
  const size_t need = al256(sizeof(float) * (size_t)n * n) + al256(sizeof(float) * (size_t)n * nr) +
                      al256(sizeof(double) * (size_t)n * nr) + al256(sizeof(double) * (size_t)n) +
                      al256(sizeof(unsigned long long) * 2 * (size_t)nr) + 256;
  char* buf = nullptr;
  int rc = mp_buffer(c, need, &buf);
  if (rc) return rc;
  char* cur = buf;
  auto take = [&](size_t bytes) { char* p = cur; cur += al256(bytes); return p; };
  float* SA = reinterpret_cast<float*>(take(sizeof(float) * (size_t)n * n));
  float* SX = reinterpret_cast<float*>(take(sizeof(float) * (size_t)n * nr));
  double* R = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n * nr));
  double* rowsum = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n));
  unsigned long long* cm = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 2 * nr));
  int* flags = reinterpret_cast<int*>(take(256));  // [0] overflow, [1] info
  const long lds = n;

  // CTE from DLANGE('I', A): row sums of |A| in column order, then the max
  // with DLANGE's NaN rule (SRC/dlange.f:173-186).
  launch_row_abssum(n, n, A, lda, rowsum, s);
  std::vector<double> rs(n);
  MP_CK(hipMemcpyAsync(rs.data(), rowsum, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  MP_CK(hipStreamSynchronize(s));
  double anrm = 0.0;
  for (int i = 0; i < n; ++i)
    if (anrm < rs[i] || std::isnan(rs[i])) anrm = rs[i];
  const double eps = 0x1p-53;  // DLAMCH('Epsilon')
  const double cte = anrm * eps * std::sqrt((double)n) * 1.0;

  auto read_flag = [&](int idx, int* v) -> int {
    MP_CK(hipMemcpyAsync(v, flags + idx, sizeof(int), hipMemcpyDeviceToHost, s));
    MP_CK(hipStreamSynchronize(s));
    return 0;
  };
  auto residual = [&]() {  // R = B - A*X (dsgesv.f:328-332)
    launch_copy_d(n, nrhs, B, ldb, R, n, s);
    launch_dgemm('N', 'N', n, nrhs, n, -1.0, A, lda, X, ldx, 1.0, R, n, s);
  };
  std::vector<unsigned long long> hn(2 * (size_t)nr);
  auto converged = [&](bool* ok) -> int {  // dsgesv.f:337-342
    *ok = true;
    if (nrhs == 0) return 0;
    launch_col_absmax(n, nrhs, X, ldx, cm, s);
    launch_col_absmax(n, nrhs, R, n, cm + nrhs, s);
    MP_CK(hipMemcpyAsync(hn.data(), cm, sizeof(unsigned long long) * 2 * nrhs, hipMemcpyDeviceToHost, s));
    MP_CK(hipStreamSynchronize(s));
    for (int i = 0; i < nrhs; ++i) {
      const double xnrm = bits_to_double(hn[i]);
      const double rnrm = bits_to_double(hn[nrhs + i]);
      if (rnrm > xnrm * cte) { *ok = false; break; }
    }
    return 0;
  };

  bool fallback = false;
  int v = 0;
  launch_fill_i32(flags, 2, 0, s);
  launch_cast_d2s(n, nrhs, B, ldb, SX, lds, flags, s);
  if ((rc = read_flag(0, &v))) return rc;
  if (v) { *iter = -2; fallback = true; }
  if (!fallback) {
    launch_cast_d2s(n, n, A, lda, SA, lds, flags, s);
    if ((rc = read_flag(0, &v))) return rc;
    if (v) { *iter = -2; fallback = true; }
  }
  if (!fallback) {
    if ((rc = getrf_device<float>(n, n, SA, lds, ipiv, flags + 1, s))) return rc;
    if ((rc = read_flag(1, &v))) return rc;
    if (v != 0) { *iter = -3; fallback = true; }
  }
  if (!fallback) {
    if (nrhs > 0 && (rc = getrs_device<float>('N', n, nrhs, SA, lds, ipiv, SX, lds, s))) return rc;
    launch_cast_s2d(n, nrhs, SX, lds, X, ldx, s);
    residual();
    bool ok = false;
    if ((rc = converged(&ok))) return rc;
    if (ok) { *iter = 0; return 0; }
    for (int it = 1; it <= itermax; ++it) {
      launch_fill_i32(flags, 1, 0, s);
      launch_cast_d2s(n, nrhs, R, n, SX, lds, flags, s);
      if ((rc = read_flag(0, &v))) return rc;
      if (v) { *iter = -2; fallback = true; break; }
      if ((rc = getrs_device<float>('N', n, nrhs, SA, lds, ipiv, SX, lds, s))) return rc;
      launch_cast_s2d(n, nrhs, SX, lds, R, n, s);
      launch_axpy(n, nrhs, R, n, X, ldx, s);
      residual();
      if ((rc = converged(&ok))) return rc;
      if (ok) { *iter = it; return 0; }
    }
    if (!fallback) *iter = -itermax - 1;
  }
  // FP64 fallback (dsgesv.f:413-427).
  launch_fill_i32(flags + 1, 1, 0, s);
  if ((rc = getrf_device<double>(n, n, A, lda, ipiv, flags + 1, s))) return rc;
  if ((rc = read_flag(1, &v))) return rc;
  if (v != 0) return v;
  launch_copy_d(n, nrhs, B, ldb, X, ldx, s);
  if (nrhs > 0 && (rc = getrs_device<double>('N', n, nrhs, A, lda, ipiv, X, ldx, s))) return rc;
  MP_CK(hipStreamSynchronize(s));
  return 0;
}

}  // namespace la
