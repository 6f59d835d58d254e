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

#include <algorithm>
#include <cmath>
#include <functional>
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
int with_safe_forms(const std::function<int()>& f);
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
  const size_t gws_n = gemv_ws_doubles(n, n, std::min(nr, 16));
  const size_t need = al256(sizeof(float) * (size_t)n * n) + al256(sizeof(float) * (size_t)n * nr) +
                      al256(sizeof(double) * (size_t)n * nr) + al256(sizeof(double) * (size_t)n) +
                      al256(sizeof(unsigned long long) * 2 * (size_t)nr) + al256(sizeof(double) * gws_n) + 256;
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
  double* gws = reinterpret_cast<double*>(take(sizeof(double) * gws_n));
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
  auto residual = [&]() {  // R = B - A*X (dsgesv.f:328-332), HBM-bound skinny products
    launch_copy_d(n, nrhs, B, ldb, R, n, s);
    for (int q0 = 0; q0 < nrhs; q0 += 16) {
      const int qn = std::min(16, nrhs - q0);
      launch_dgemv_n(n, n, qn, -1.0, A, lda, X + (long)q0 * ldx, ldx, 1.0, R + (long)q0 * n, n, gws, s);
    }
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

// ============================================================================
// Equilibration (DGEEQU / DGEEQUB, SRC/dgeequ.f:175-296, SRC/dgeequb.f:182-313)
// ============================================================================
namespace la {

// Row and column scale factors of an m x n device matrix.  The maxima are
// taken on the GPU (exact: max of absolute values; the column pass's product
// |a(i,j)|*r(i) is the reference's single multiply); the decisions, the
// radix rounding (radix != 0: DGEEQUB's powers of two) and the reciprocals
// run on the host with the reference's formulas.  r (m) and c (n) are host
// arrays; d_r, d_c device arrays of the same sizes receive copies.
int geequ_device(int m, int n, const double* A, long lda, double* r, double* c, double* rowcnd, double* colcnd,
                 double* amax, int radix, double* d_r, double* d_c, hipStream_t s) {
  if (m == 0 || n == 0) {
    *rowcnd = 1.0;
    *colcnd = 1.0;
    *amax = 0.0;
    return 0;
  }
  const double smlnum = 0x1p-1022;  // DLAMCH('S')
  const double bignum = 1.0 / smlnum;
  const double logrdx = std::log(2.0);
  launch_row_absmax(m, n, A, lda, d_r, s);
  MP_CK(hipMemcpyAsync(r, d_r, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  MP_CK(hipStreamSynchronize(s));
  if (radix)
    for (int i = 0; i < m; ++i)
      if (r[i] > 0) r[i] = std::pow(2.0, (int)(std::log(r[i]) / logrdx));
  double rcmin = bignum, rcmax = 0.0;
  for (int i = 0; i < m; ++i) {
    rcmax = std::max(rcmax, r[i]);
    rcmin = std::min(rcmin, r[i]);
  }
  *amax = rcmax;
  if (rcmin == 0.0) {
    for (int i = 0; i < m; ++i)
      if (r[i] == 0.0) return i + 1;
  }
  for (int i = 0; i < m; ++i) r[i] = 1.0 / std::min(std::max(r[i], smlnum), bignum);
  *rowcnd = std::max(rcmin, smlnum) / std::min(rcmax, bignum);
  MP_CK(hipMemcpyAsync(d_r, r, sizeof(double) * m, hipMemcpyHostToDevice, s));
  launch_col_absmax_scaled(m, n, A, lda, d_r, d_c, s);
  MP_CK(hipMemcpyAsync(c, d_c, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  MP_CK(hipStreamSynchronize(s));
  for (int j = 0; j < n; ++j)
    if (radix && c[j] > 0) c[j] = std::pow(2.0, (int)(std::log(c[j]) / logrdx));
  rcmin = bignum;
  rcmax = 0.0;
  for (int j = 0; j < n; ++j) {
    rcmin = std::min(rcmin, c[j]);
    rcmax = std::max(rcmax, c[j]);
  }
  if (rcmin == 0.0) {
    for (int j = 0; j < n; ++j)
      if (c[j] == 0.0) return m + j + 1;
  }
  for (int j = 0; j < n; ++j) c[j] = 1.0 / std::min(std::max(c[j], smlnum), bignum);
  *colcnd = std::max(rcmin, smlnum) / std::min(rcmax, bignum);
  MP_CK(hipMemcpyAsync(d_c, c, sizeof(double) * n, hipMemcpyHostToDevice, s));
  MP_CK(hipStreamSynchronize(s));
  return 0;
}

// DLAQGE (SRC/dlaqge.f:183-226): decide EQUED and scale the device matrix.
char laqge_device(int m, int n, double* A, long lda, const double* d_r, const double* d_c, double rowcnd,
                  double colcnd, double amax, hipStream_t s) {
  if (m <= 0 || n <= 0) return 'N';
  const double thresh = 0.1;
  const double small = 0x1p-1022 / 0x1p-52;  // DLAMCH('S') / DLAMCH('P')
  const double large = 1.0 / small;
  int mode;
  char equed;
  if (rowcnd >= thresh && amax >= small && amax <= large) {
    if (colcnd >= thresh) return 'N';
    mode = 2;
    equed = 'C';
  } else if (colcnd >= thresh) {
    mode = 1;
    equed = 'R';
  } else {
    mode = 3;
    equed = 'B';
  }
  launch_scale_rc(m, n, A, lda, d_r, d_c, mode, s);
  return equed;
}

// ============================================================================
// Mixed-precision solve with refinement and GMRES-IR
// ============================================================================
int getrf_device_lp(int m, int n, float* A, long lda, int* ipiv, int* info, int kind, hipStream_t s);

namespace {

struct IrBufs {
  float* SA;
  float* SX;   // n x nr
  double* R;   // n x nr
  double* rr;  // row scales (device)
  double* cc;  // column scales (device)
  double* V;   // n x (m+1) GMRES basis
  double* t;   // n
  double* w;   // n
  double* h;   // 2*(m+2): h1, h2
  double* sc;  // scalars
  double* gws; // gemv workspace
  double* LU64; // FP64 copy of the low-precision factors (GMRES's preconditioner)
  double* Minv; // C (LU)^-1 R formed in FP64 from LU64 (n <= 16384), or null
  unsigned long long* cm;
  int* flags;
};

}  // namespace

// Solve A X = B to the FP64 stopping test of DSGESV (max|r| <= max|x| * ||A||_inf
// * eps * sqrt(n) per column, SRC/dsgesv.f:286-288,337-342) with an LU
// factored in low precision: kind 0 FP32, 1 FP16 updates, 2 BF16 updates.
// For kinds 1 and 2 the matrix is first equilibrated by DGEEQUB's powers of
// two (exact), so the factored matrix is R*A*C and corrections are
// C * (LU)^-1 * R * r.  Classic refinement (corrections solved in FP32)
// runs while the worst backward error drops at least tenfold per step; then
// restarted GMRES(50) on A d = r, left-preconditioned by the low-precision
// LU applied in FP64 (GMRES-IR), FP64 Gram-Schmidt (two passes); then, if
// still unconverged, the FP64 LU (as DSGESV's fallback).  *iter: refinement
// steps (classic + GMRES cycles) on success, -2 (overflow in the cast), -3
// (singular low-precision LU) or -31 (no convergence) when the FP64 fallback
// produced X.  *gmres_iter: GMRES inner iterations used.  berr (nrhs
// doubles, may be null): final normwise backward errors
// max|r|/(||A||_inf max|x|); NaN after the FP64 fallback (A then holds the
// factors).
int dgesv_ir_device(int n, int nrhs, double* A, long lda, int* ipiv, const double* B, long ldb, double* X, long ldx,
                    int kind, int* iter, int* gmres_iter, double* berr, hipStream_t s) {
  *iter = 0;
  *gmres_iter = 0;
  if (n == 0) return 0;
  DevCtx* ctx = ctx_for_current();
  if (!ctx) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(mp_mutex(ctx));
  // GMRES basis: up to min(n, 8192) vectors (restart length), within about
  // 2 GiB; the explicit inverse of the low-precision LU (GMRES's
  // preconditioner, formed once in FP64) for n <= 16384
  const int GM = (int)std::max<long>(50, std::min<long>({(long)n, 8192L, (2L << 30) / (8L * n)}));
  const bool minv_form = n <= 16384;
  const int nr = std::max(nrhs, 1);
  const int rch = std::min(nr, 16);
  const size_t gws_n = gemv_ws_doubles(n, std::max(n, GM + 1), rch);
  const size_t need = al256(sizeof(float) * (size_t)n * n) + al256(sizeof(double) * (size_t)n * n) +
                      (minv_form ? al256(sizeof(double) * (size_t)n * n) : 0) +
                      al256(sizeof(float) * (size_t)n * nr) +
                      al256(sizeof(double) * (size_t)n * nr) + 2 * al256(sizeof(double) * (size_t)n) +
                      al256(sizeof(double) * (size_t)n * (GM + 1)) + 2 * al256(sizeof(double) * (size_t)n) +
                      al256(sizeof(double) * (2 * (size_t)GM + 8)) + al256(sizeof(double) * 64) +
                      al256(sizeof(double) * gws_n) +
                      al256(sizeof(unsigned long long) * 2 * (size_t)nr) + 256;
  char* buf = nullptr;
  int rc = mp_buffer(ctx, need, &buf);
  if (rc) return rc;
  char* cur = buf;
  auto take = [&](size_t bytes) { char* p = cur; cur += al256(bytes); return p; };
  IrBufs b;
  b.SA = reinterpret_cast<float*>(take(sizeof(float) * (size_t)n * n));
  b.LU64 = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n * n));
  b.Minv = minv_form ? reinterpret_cast<double*>(take(sizeof(double) * (size_t)n * n)) : nullptr;
  b.SX = reinterpret_cast<float*>(take(sizeof(float) * (size_t)n * nr));
  b.R = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n * nr));
  b.rr = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n));
  b.cc = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n));
  b.V = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n * (GM + 1)));
  b.t = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n));
  b.w = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n));
  b.h = reinterpret_cast<double*>(take(sizeof(double) * (2 * (size_t)GM + 8)));
  b.sc = reinterpret_cast<double*>(take(sizeof(double) * 64));
  b.gws = reinterpret_cast<double*>(take(sizeof(double) * gws_n));
  b.cm = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 2 * nr));
  b.flags = reinterpret_cast<int*>(take(256));
  const long ln = n;

  // ||A||_inf and the stopping constant
  launch_row_abssum(n, n, A, lda, b.t, s);
  std::vector<double> hv(n);
  MP_CK(hipMemcpyAsync(hv.data(), b.t, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  MP_CK(hipStreamSynchronize(s));
  double anrm = 0.0;
  for (int i = 0; i < n; ++i)
    if (anrm < hv[i] || std::isnan(hv[i])) anrm = hv[i];
  const double eps = 0x1p-53;
  const double cte = anrm * eps * std::sqrt((double)n);

  // equilibration for the half-precision factorizations
  const double* rr = nullptr;
  const double* cc = nullptr;
  if (kind == 1 || kind == 2) {
    std::vector<double> hr(n), hc(n);
    double rowcnd, colcnd, amax;
    const int einfo = geequ_device(n, n, A, lda, hr.data(), hc.data(), &rowcnd, &colcnd, &amax, 1, b.rr, b.cc, s);
    if (einfo < 0) return einfo;
    if (einfo == 0) {
      rr = b.rr;
      cc = b.cc;
    }
  }

  auto read_flag = [&](int idx, int* v) -> int {
    MP_CK(hipMemcpyAsync(v, b.flags + idx, sizeof(int), hipMemcpyDeviceToHost, s));
    MP_CK(hipStreamSynchronize(s));
    return 0;
  };
  // y <- C * (LU)^-1 * R * v for nrhs columns (add: y += ...)
  auto precond = [&](int cols, const double* v, long ldv, double* y, long ldy, int add) -> int {
    launch_scale_cast_d2s(n, cols, v, ldv, rr, nullptr, b.SX, ln, nullptr, s);
    int r2 = getrs_device<float>('N', n, cols, b.SA, ln, ipiv, b.SX, ln, s);
    if (r2) return r2;
    launch_scale_s2d(n, cols, b.SX, ln, cc, y, ldy, add, s);
    return 0;
  };
  // GMRES's preconditioner, applied in FP64 with the low-precision factors
  // (the working-precision application of GMRES-IR): y <- C (LU)^-1 R v.
  // For n <= 16384 the operator is formed once, explicitly: (LU)^-1 by the
  // blocked triangular solves on the identity (no waits, MFMA products),
  // then scaled by C and R, so each application is one pass over an n x n
  // matrix instead of two latency-bound sweeps.
  bool lu64_ready = false;
  auto precond64 = [&](const double* v, double* y) -> int {
    if (!lu64_ready) {
      launch_cast_s2d(n, n, b.SA, ln, b.LU64, ln, s);
      if (b.Minv) {
        launch_identity(n, b.Minv, ln, s);
        int r3 = with_safe_forms([&]() { return getrs_device<double>('N', n, n, b.LU64, ln, ipiv, b.Minv, ln, s); });
        if (r3) return r3;
        if (rr || cc) launch_scale_rc(n, n, b.Minv, ln, cc, rr, cc && rr ? 3 : cc ? 1 : 2, s);
      }
      lu64_ready = true;
    }
    if (b.Minv) {
      launch_dgemv_n(n, n, 1, 1.0, b.Minv, ln, v, ln, 0.0, y, ln, b.gws, s);
      return 0;
    }
    if (v != y) launch_copy_d(n, 1, v, ln, y, ln, s);
    if (rr) launch_scale_rc(n, 1, y, ln, rr, nullptr, 1, s);
    int r2 = getrs_device<double>('N', n, 1, b.LU64, ln, ipiv, y, ln, s);
    if (r2) return r2;
    if (cc) launch_scale_rc(n, 1, y, ln, cc, nullptr, 1, s);
    return 0;
  };
  auto residual = [&]() {  // R = B - A*X
    launch_copy_d(n, nrhs, B, ldb, b.R, ln, s);
    for (int q0 = 0; q0 < nrhs; q0 += 16) {
      const int qn = std::min(16, nrhs - q0);
      launch_dgemv_n(n, n, qn, -1.0, A, lda, X + (long)q0 * ldx, ldx, 1.0, b.R + (long)q0 * ln, ln, b.gws, s);
    }
  };
  std::vector<unsigned long long> hn(2 * (size_t)nr);
  std::vector<double> be(nr, 0.0);
  std::vector<char> done(nr, 0);
  auto check = [&](bool* all) -> int {  // DSGESV's per-column test; be = backward errors
    *all = true;
    launch_col_absmax(n, nrhs, X, ldx, b.cm, s);
    launch_col_absmax(n, nrhs, b.R, ln, b.cm + nrhs, s);
    MP_CK(hipMemcpyAsync(hn.data(), b.cm, sizeof(unsigned long long) * 2 * nrhs, hipMemcpyDeviceToHost, s));
    MP_CK(hipStreamSynchronize(s));
    for (int i = 0; i < nrhs; ++i) {
      const double xnrm = bits_to_double(hn[i]);
      const double rnrm = bits_to_double(hn[nrhs + i]);
      be[i] = rnrm / (anrm * xnrm);
      done[i] = !(rnrm > xnrm * cte);
      if (!done[i]) *all = false;
    }
    return 0;
  };

  bool fallback = false;
  int v = 0;
  launch_fill_i32(b.flags, 2, 0, s);
  launch_scale_cast_d2s(n, n, A, lda, rr, cc, b.SA, ln, b.flags, s);
  if ((rc = read_flag(0, &v))) return rc;
  if (v) { *iter = -2; fallback = true; }
  if (!fallback) {
    if ((rc = getrf_device_lp(n, n, b.SA, ln, ipiv, b.flags + 1, kind, s))) return rc;
    if ((rc = read_flag(1, &v))) return rc;
    if (v != 0) { *iter = -3; fallback = true; }
  }
  if (!fallback && nrhs > 0) {
    for (int q0 = 0; q0 < nrhs; q0 += 16) {
      const int qn = std::min(16, nrhs - q0);
      if ((rc = precond(qn, B + (long)q0 * ldb, ldb, X + (long)q0 * ldx, ldx, 0))) return rc;
    }
    residual();
    bool all = false;
    if ((rc = check(&all))) return rc;
    int steps = 0;
    double prev = 1e300;
    const int max_classic = 30, max_cycles = 30;
    // classic refinement while the worst backward error at least halves
    while (!all && steps < max_classic) {
      double worst = 0.0;
      for (int i = 0; i < nrhs; ++i) worst = std::max(worst, be[i]);
      if (steps > 0 && !(worst < 0.1 * prev)) break;  // slower than 10x per step: GMRES-IR
      prev = worst;
      for (int q0 = 0; q0 < nrhs; q0 += 16) {
        const int qn = std::min(16, nrhs - q0);
        if ((rc = precond(qn, b.R + (long)q0 * ln, ln, X + (long)q0 * ldx, ldx, 1))) return rc;
      }
      ++steps;
      residual();
      if ((rc = check(&all))) return rc;
    }
    // GMRES-IR on the columns still unconverged: GMRES(GM) on A d = r,
    // left-preconditioned, classical Gram-Schmidt applied twice in FP64,
    // Givens rotations and the small triangular solve on the host.  At most
    // 30 restarts and 2n inner iterations in all.
    int cycles = 0;
    const long max_inner = 2L * n;
    // Hessenberg columns after rotation (column j holds j + 1 entries at
    // offset j (j + 1) / 2), Givens cosines / sines, the rotated residual
    std::vector<double> H, cs(GM, 0.0), sn(GM, 0.0), gv(GM + 1, 0.0), hh(2 * (size_t)GM + 8);
    while (!all && cycles < max_cycles && *gmres_iter < max_inner) {
      for (int col = 0; col < nrhs; ++col) {
        if (done[col]) continue;
        double* xc = X + (long)col * ldx;
        // z = M^-1 r, beta = ||z||_2, V0 = z / beta
        if ((rc = precond64(b.R + (long)col * ln, b.w))) return rc;
        launch_dgemv_cols(n, 1, 1.0, b.w, ln, b.w, 0.0, b.sc, false, s);
        double nrm2 = 0.0;
        MP_CK(hipMemcpyAsync(&nrm2, b.sc, sizeof(double), hipMemcpyDeviceToHost, s));
        MP_CK(hipStreamSynchronize(s));
        const double beta = std::sqrt(nrm2);
        if (!(beta > 0.0) || !std::isfinite(beta)) continue;
        H.assign((size_t)GM * (GM + 1) / 2 + GM, 0.0);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        {
          const double inv = 1.0 / beta;
          MP_CK(hipMemcpyAsync(b.sc + 1, &inv, sizeof(double), hipMemcpyHostToDevice, s));
          // V0 = w * (1/beta) via the product of a 1-column matrix with a 1-vector
          launch_dgemv_n(n, 1, 1, 1.0, b.w, ln, b.sc + 1, 1, 0.0, b.V, ln, b.gws, s);
        }
        int k = 0;
        const double tol = 1e-12;
        for (; k < GM; ++k) {
          double* vk = b.V + (long)k * ln;
          const int kk = k + 1;  // basis vectors so far
          launch_dgemv_n(n, n, 1, 1.0, A, lda, vk, ln, 0.0, b.t, ln, b.gws, s);   // t = A v_k
          if ((rc = precond64(b.t, b.w))) return rc;                              // w = M^-1 t
          // two passes of classical Gram-Schmidt: h1 = V^T w, w -= V h1;
          // h2 = V^T w, w -= V h2; then ||w||^2 -- one copy to the host
          launch_dgemv_cols(n, kk, 1.0, b.V, ln, b.w, 0.0, b.h, false, s);
          launch_dgemv_n(n, kk, 1, -1.0, b.V, ln, b.h, kk, 1.0, b.w, ln, b.gws, s);
          launch_dgemv_cols(n, kk, 1.0, b.V, ln, b.w, 0.0, b.h + kk, false, s);
          launch_dgemv_n(n, kk, 1, -1.0, b.V, ln, b.h + kk, kk, 1.0, b.w, ln, b.gws, s);
          launch_dgemv_cols(n, 1, 1.0, b.w, ln, b.w, 0.0, b.h + 2 * kk, false, s);
          MP_CK(hipMemcpyAsync(hh.data(), b.h, sizeof(double) * (2 * (size_t)kk + 1), hipMemcpyDeviceToHost, s));
          MP_CK(hipStreamSynchronize(s));
          ++*gmres_iter;
          double* hc = &H[(size_t)k * (k + 1) / 2];
          std::vector<double> col_h(kk + 1);
          for (int i = 0; i < kk; ++i) col_h[i] = hh[i] + hh[kk + i];
          const double hn1 = std::sqrt(hh[2 * kk]);
          col_h[kk] = hn1;
          for (int i = 0; i < k; ++i) {  // previous rotations
            const double t1 = cs[i] * col_h[i] + sn[i] * col_h[i + 1];
            col_h[i + 1] = -sn[i] * col_h[i] + cs[i] * col_h[i + 1];
            col_h[i] = t1;
          }
          const double den = std::hypot(col_h[k], col_h[kk]);
          cs[k] = den > 0 ? col_h[k] / den : 1.0;
          sn[k] = den > 0 ? col_h[kk] / den : 0.0;
          col_h[k] = den;
          for (int i = 0; i < kk; ++i) hc[i] = col_h[i];
          gv[kk] = -sn[k] * gv[k];
          gv[k] = cs[k] * gv[k];
          const bool stop = std::fabs(gv[kk]) <= tol * beta || !(hn1 > 0.0) || kk == GM ||
                            *gmres_iter >= max_inner;
          if (!stop) {
            const double inv = 1.0 / hn1;
            MP_CK(hipMemcpyAsync(b.sc + 1, &inv, sizeof(double), hipMemcpyHostToDevice, s));
            launch_dgemv_n(n, 1, 1, 1.0, b.w, ln, b.sc + 1, 1, 0.0, b.V + (long)kk * ln, ln, b.gws, s);
          } else {
            k = kk;
            break;
          }
        }
        // y = Rk^-1 g (Rk the rotated Hessenberg, upper triangular), x += V y
        std::vector<double> y(k, 0.0);
        for (int i = k - 1; i >= 0; --i) {
          double sacc = gv[i];
          for (int j = i + 1; j < k; ++j) sacc -= H[(size_t)j * (j + 1) / 2 + i] * y[j];
          y[i] = sacc / H[(size_t)i * (i + 1) / 2 + i];
        }
        MP_CK(hipMemcpyAsync(b.h, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        launch_dgemv_n(n, k, 1, 1.0, b.V, ln, b.h, k, 1.0, xc, ldx, b.gws, s);
        MP_CK(hipStreamSynchronize(s));  // y's host buffer is released below
      }
      ++cycles;
      residual();
      if ((rc = check(&all))) return rc;
    }
    *iter = steps + cycles;
    if (berr)
      for (int i = 0; i < nrhs; ++i) berr[i] = be[i];
    if (all) {
      MP_CK(hipStreamSynchronize(s));
      return 0;
    }
    *iter = -31;
  } else if (!fallback) {
    *iter = 0;  // nrhs == 0
    MP_CK(hipStreamSynchronize(s));
    return 0;
  }
  // FP64 fallback (as SRC/dsgesv.f:413-427)
  launch_fill_i32(b.flags + 1, 1, 0, s);
  if ((rc = getrf_device<double>(n, n, A, lda, ipiv, b.flags + 1, s))) return rc;
  if ((rc = read_flag(1, &v))) return rc;
  if (v != 0) return v;
  launch_copy_d(n, nrhs, B, ldb, X, ldx, s);
  if (nrhs > 0 && (rc = getrs_device<double>('N', n, nrhs, A, lda, ipiv, X, ldx, s))) return rc;
  // A now holds the FP64 factors: the backward error is not recomputed.
  if (berr)
    for (int i = 0; i < nrhs; ++i) berr[i] = std::nan("");
  MP_CK(hipStreamSynchronize(s));
  return 0;
}

}  // namespace la

// ============================================================================
// DGERFS: FP64 iterative refinement with error bounds (SRC/dgerfs.f:262-432)
// ============================================================================
namespace la {

void lacn2_host(int n, double* v, double* x, int* isgn, double* est, int* kase, int* isave);

// All matrices are device memory; ferr/berr host arrays.  The O(n^2) work
// (residual, |op(A)||x|, the solves with the factors) runs on the GPU; the
// reference's scalar decisions (stopping test, safe divisions, DLACN2's
// reverse communication) run on the host on n-vectors.
int gerfs_device(char trans, int n, int nrhs, const double* A, long lda, const double* AF, long ldaf,
                 const int* ipiv, const double* B, long ldb, double* X, long ldx, double* ferr, double* berr,
                 hipStream_t s) {
  if (n == 0 || nrhs == 0) {
    for (int j = 0; j < nrhs; ++j) { ferr[j] = 0.0; berr[j] = 0.0; }
    return 0;
  }
  DevCtx* ctx = ctx_for_current();
  if (!ctx) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(mp_mutex(ctx));
  const bool notran = (trans == 'N' || trans == 'n');
  const char transt = notran ? 'T' : 'N';
  const size_t gws_n = gemv_ws_doubles(n, n, 1);
  char* buf = nullptr;
  int rc = mp_buffer(ctx, 2 * al256(sizeof(double) * (size_t)n) + al256(sizeof(double) * gws_n), &buf);
  if (rc) return rc;
  double* d_w1 = reinterpret_cast<double*>(buf);
  double* d_w2 = reinterpret_cast<double*>(buf + al256(sizeof(double) * (size_t)n));
  double* gws = reinterpret_cast<double*>(buf + 2 * al256(sizeof(double) * (size_t)n));
  const int itmax = 5;
  const int nz = n + 1;
  const double eps = 0x1p-53, safmin = 0x1p-1022;
  const double safe1 = nz * safmin, safe2 = safe1 / eps;
  std::vector<double> w1(n), w2(n), bj(n), xv(n), v(n);
  std::vector<int> isgn(n);
  for (int j = 0; j < nrhs; ++j) {
    double* xj = X + (long)j * ldx;
    const double* bjd = B + (long)j * ldb;
    MP_CK(hipMemcpyAsync(bj.data(), bjd, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    int count = 1;
    double lstres = 3.0;
    while (true) {
      // w2 = b - op(A) x; w1 = |op(A)| |x|
      launch_copy_d(n, 1, bjd, ldb, d_w2, n, s);
      if (notran) {
        launch_dgemv_n(n, n, 1, -1.0, A, lda, xj, ldx, 1.0, d_w2, n, gws, s);
        launch_dgemv_n_abs(n, n, 1.0, A, lda, xj, 0.0, d_w1, gws, s);
      } else {
        launch_dgemv_cols(n, n, -1.0, A, lda, xj, 1.0, d_w2, false, s);
        launch_dgemv_cols(n, n, 1.0, A, lda, xj, 0.0, d_w1, true, s);
      }
      MP_CK(hipMemcpyAsync(w1.data(), d_w1, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      MP_CK(hipMemcpyAsync(w2.data(), d_w2, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      MP_CK(hipStreamSynchronize(s));
      double sb = 0.0;
      for (int i = 0; i < n; ++i) {
        w1[i] = std::fabs(bj[i]) + w1[i];
        if (w1[i] > safe2) sb = std::max(sb, std::fabs(w2[i]) / w1[i]);
        else sb = std::max(sb, (std::fabs(w2[i]) + safe1) / (w1[i] + safe1));
      }
      berr[j] = sb;
      if (berr[j] > eps && 2.0 * berr[j] <= lstres && count <= itmax) {
        if ((rc = getrs_device<double>(trans, n, 1, AF, ldaf, ipiv, d_w2, n, s))) return rc;
        launch_axpy(n, 1, d_w2, n, xj, ldx, s);
        lstres = berr[j];
        ++count;
        continue;
      }
      break;
    }
    // FERR by DLACN2 on inv(op(A)) * diag(W)
    for (int i = 0; i < n; ++i) {
      if (w1[i] > safe2) w1[i] = std::fabs(w2[i]) + nz * eps * w1[i];
      else w1[i] = std::fabs(w2[i]) + nz * eps * w1[i] + safe1;
    }
    int kase = 0;
    int isave[3] = {0, 0, 0};
    double est = 0.0;
    std::vector<double> xw(n);
    while (true) {
      lacn2_host(n, v.data(), xw.data(), isgn.data(), &est, &kase, isave);
      if (kase == 0) break;
      if (kase == 1) {
        MP_CK(hipMemcpyAsync(d_w2, xw.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        if ((rc = getrs_device<double>(transt, n, 1, AF, ldaf, ipiv, d_w2, n, s))) return rc;
        MP_CK(hipMemcpyAsync(xw.data(), d_w2, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        MP_CK(hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i) xw[i] = w1[i] * xw[i];
      } else {
        for (int i = 0; i < n; ++i) xw[i] = w1[i] * xw[i];
        MP_CK(hipMemcpyAsync(d_w2, xw.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        if ((rc = getrs_device<double>(trans, n, 1, AF, ldaf, ipiv, d_w2, n, s))) return rc;
        MP_CK(hipMemcpyAsync(xw.data(), d_w2, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        MP_CK(hipStreamSynchronize(s));
      }
    }
    ferr[j] = est;
    MP_CK(hipMemcpyAsync(xv.data(), xj, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    MP_CK(hipStreamSynchronize(s));
    double xmax = 0.0;
    for (int i = 0; i < n; ++i) xmax = std::max(xmax, std::fabs(xv[i]));
    if (xmax != 0.0) ferr[j] = ferr[j] / xmax;
  }
  return 0;
}

// DLACN2 (SRC/dlacn2.f:167-282): the 1-norm estimator's reverse
// communication, on host vectors.
void lacn2_host(int n, double* v, double* x, int* isgn, double* est, int* kase, int* isave) {
  const int itmax = 5;
  auto idamax = [&](const double* y) {  // first index of max |y| (BLAS/SRC/idamax.f)
    int im = 1;
    double dm = std::fabs(y[0]);
    for (int i = 2; i <= n; ++i)
      if (std::fabs(y[i - 1]) > dm) { im = i; dm = std::fabs(y[i - 1]); }
    return im;
  };
  auto asum = [&](const double* y) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t = t + std::fabs(y[i]);
    return t;
  };
  auto restart = [&]() {  // DLACN2 label 50
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    x[isave[1] - 1] = 1.0;
    *kase = 1;
    isave[0] = 3;
  };
  auto final_stage = [&]() {  // label 120
    double alt = 1.0;
    for (int i = 1; i <= n; ++i) {
      x[i - 1] = alt * (1.0 + (double)(i - 1) / (double)(n - 1));
      alt = -alt;
    }
    *kase = 1;
    isave[0] = 5;
  };
  if (*kase == 0) {
    for (int i = 0; i < n; ++i) x[i] = 1.0 / (double)n;
    *kase = 1;
    isave[0] = 1;
    return;
  }
  switch (isave[0]) {
    case 1:
      if (n == 1) {
        v[0] = x[0];
        *est = std::fabs(v[0]);
        *kase = 0;
        return;
      }
      *est = asum(x);
      for (int i = 0; i < n; ++i) {
        x[i] = x[i] >= 0.0 ? 1.0 : -1.0;
        isgn[i] = (int)std::lround(x[i]);
      }
      *kase = 2;
      isave[0] = 2;
      return;
    case 2:
      isave[1] = idamax(x);
      isave[2] = 2;
      restart();
      return;
    case 3: {
      for (int i = 0; i < n; ++i) v[i] = x[i];
      const double estold = *est;
      *est = asum(v);
      bool same = true;
      for (int i = 0; i < n; ++i)
        if ((int)std::lround(x[i] >= 0.0 ? 1.0 : -1.0) != isgn[i]) { same = false; break; }
      if (same || *est <= estold) { final_stage(); return; }
      for (int i = 0; i < n; ++i) {
        x[i] = x[i] >= 0.0 ? 1.0 : -1.0;
        isgn[i] = (int)std::lround(x[i]);
      }
      *kase = 2;
      isave[0] = 4;
      return;
    }
    case 4: {
      const int jlast = isave[1];
      isave[1] = idamax(x);
      if (x[jlast - 1] != std::fabs(x[isave[1] - 1]) && isave[2] < itmax) {
        isave[2] += 1;
        restart();
        return;
      }
      final_stage();
      return;
    }
    case 5: {
      const double temp = 2.0 * (asum(x) / (double)(3 * n));
      if (temp > *est) {
        for (int i = 0; i < n; ++i) v[i] = x[i];
        *est = temp;
      }
      *kase = 0;
      return;
    }
  }
}


// ============================================================================
// DGECON: reciprocal condition number estimate from the factors of DGETRF
// (SRC/dgecon.f:130-244)
// ============================================================================
// DLACN2's reverse communication on the host (n-vectors), the triangular
// solves with L and U (or their transposes) on the GPU, one tri_chain sweep
// each.  DLATRS's scaling (SRC/dlatrs.f) only departs from a plain solve when
// the solution would overflow; there the reference finds SCALE < |x|*SMLNUM
// and returns RCOND = 0 (dgecon.f:226-229), which is what a solve that
// overflows to a non-finite value returns here.  anorm: the norm of the
// original matrix, checked by the caller.  Returns INFO (0, or 1 when the
// estimate of ||inv(A)|| is zero or RCOND is NaN / above the overflow
// threshold, dgecon.f:233-242) or a negative LA_ERR_* code.
int gecon_device(bool onenrm, int n, const double* AF, long ldaf, double anorm, double* rcond, hipStream_t s) {
  *rcond = 0.0;
  if (n == 0) {
    *rcond = 1.0;
    return 0;
  }
  if (anorm == 0.0) return 0;
  DevCtx* ctx = ctx_for_current();
  if (!ctx) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(mp_mutex(ctx));
  char* buf = nullptr;
  const size_t xb = al256(sizeof(double) * (size_t)n);
  int rc = mp_buffer(ctx, 2 * xb + al256(sizeof(unsigned) * tri_chain_ctl_words(1)) + 256, &buf);
  if (rc) return rc;
  double* d_x = reinterpret_cast<double*>(buf);
  double* d_y = reinterpret_cast<double*>(buf + xb);
  unsigned* ctl = reinterpret_cast<unsigned*>(buf + 2 * xb);
  int* err = reinterpret_cast<int*>(buf + 2 * xb + al256(sizeof(unsigned) * tri_chain_ctl_words(1)));
  const int kase1 = onenrm ? 1 : 2;
  std::vector<double> v(n), x(n);
  std::vector<int> isgn(n);
  int kase = 0, isave[3] = {0, 0, 0};
  double ainvnm = 0.0;
  launch_fill_i32(err, 1, 0, s);
  while (true) {
    lacn2_host(n, v.data(), x.data(), isgn.data(), &ainvnm, &kase, isave);
    if (kase == 0) break;
    MP_CK(hipMemcpyAsync(d_x, x.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (kase == kase1) {
      launch_tri_chain<double>(0, n, 1, AF, ldaf, d_x, n, d_y, n, ctl, err, s);  // inv(L)
      launch_tri_chain<double>(1, n, 1, AF, ldaf, d_y, n, d_x, n, ctl, err, s);  // inv(U)
    } else {
      launch_tri_chain<double>(2, n, 1, AF, ldaf, d_x, n, d_y, n, ctl, err, s);  // inv(U^T)
      launch_tri_chain<double>(3, n, 1, AF, ldaf, d_y, n, d_x, n, ctl, err, s);  // inv(L^T)
    }
    int e = 0;
    MP_CK(hipMemcpyAsync(x.data(), d_x, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    MP_CK(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, s));
    MP_CK(hipStreamSynchronize(s));
    if (e) return LA_ERR_SYNC;
    for (int i = 0; i < n; ++i)
      if (!std::isfinite(x[i])) return 0;  // DLATRS would scale to zero: RCOND = 0
  }
  if (ainvnm != 0.0) {
    *rcond = (1.0 / ainvnm) / anorm;
  } else {
    return 1;
  }
  if (std::isnan(*rcond) || *rcond > 1.79769313486231571e308) return 1;
  return 0;
}

// ============================================================================
// DGETRI: inverse from the factors of DGETRF (SRC/dgetri.f:114-263)
// ============================================================================
// INFO = i when U(i,i) is exactly zero (dgetri.f:150-155), A untouched then.
// Otherwise inv(A) = inv(U) inv(L) P is formed as the solution of
// A X = I with the saved factors -- the blocked triangular solves, whose
// products run on the FP64 matrix cores -- in a device buffer, then copied
// over A.  The reference forms inv(U) in place and solves inv(A) L = inv(U);
// the results agree to rounding (DGET03's ratio, TESTING/LIN/dget03.f).
int getri_device(int n, double* A, long lda, const int* ipiv, hipStream_t s) {
  if (n == 0) return 0;
  DevCtx* ctx = ctx_for_current();
  if (!ctx) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(mp_mutex(ctx));
  std::vector<double> diag(n);
  MP_CK(hipMemcpy2DAsync(diag.data(), sizeof(double), A, sizeof(double) * (size_t)(lda + 1), sizeof(double), n,
                         hipMemcpyDeviceToHost, s));
  MP_CK(hipStreamSynchronize(s));
  for (int i = 0; i < n; ++i)
    if (diag[i] == 0.0) return i + 1;
  char* buf = nullptr;
  int rc = mp_buffer(ctx, al256(sizeof(double) * (size_t)n * n), &buf);
  if (rc) return rc;
  double* W = reinterpret_cast<double*>(buf);
  launch_identity(n, W, n, s);
  rc = with_safe_forms([&]() { return getrs_device<double>('N', n, n, A, lda, ipiv, W, n, s); });
  if (rc) return rc;
  launch_copy_d(n, n, W, n, A, lda, s);
  MP_CK(hipStreamSynchronize(s));
  return 0;
}


// ============================================================================
// DLANGE on a device matrix (SRC/dlange.f:143-200)
// ============================================================================
// 'M' max |a|, '1'/'O' max column sum, 'I' max row sum (sums in the
// reference's order, so bitwise its results), 'F'/'E' Frobenius norm as
// scale * sqrt(sum (a/scale)^2) with scale = max |a| (the reference's DLASSQ
// accumulates with other scalings: equal to rounding).  The maximum over
// columns / rows follows the reference's NaN rule (a NaN wins).
int lange_device(char norm, int m, int n, const double* A, long lda, double* value, hipStream_t s) {
  *value = 0.0;
  if (std::min(m, n) == 0) return 0;
  DevCtx* ctx = ctx_for_current();
  if (!ctx) return LA_ERR_NODEVICE;
  std::lock_guard<std::mutex> g(mp_mutex(ctx));
  const int len = std::max(m, n);
  char* buf = nullptr;
  int rc = mp_buffer(ctx, al256(sizeof(double) * (size_t)len), &buf);
  if (rc) return rc;
  double* d_v = reinterpret_cast<double*>(buf);
  std::vector<double> h(len);
  auto fold_max = [&](int cnt) {
    double v = 0.0;
    for (int i = 0; i < cnt; ++i)
      if (v < h[i] || std::isnan(h[i])) v = h[i];
    return v;
  };
  auto fetch = [&](int cnt) -> int {
    MP_CK(hipMemcpyAsync(h.data(), d_v, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
    MP_CK(hipStreamSynchronize(s));
    return 0;
  };
  const char c = (char)(norm | 0x20);
  if (c == 'm' || c == 'f' || c == 'e') {
    launch_col_norm(m, n, A, lda, 0, 1.0, d_v, s);
    if ((rc = fetch(n))) return rc;
    const double amax = fold_max(n);
    if (c == 'm' || !(amax > 0.0) || !std::isfinite(amax)) {
      *value = amax;
      return 0;
    }
    launch_col_norm(m, n, A, lda, 2, amax, d_v, s);
    if ((rc = fetch(n))) return rc;
    double ssq = 0.0;
    for (int j = 0; j < n; ++j) ssq += h[j];
    *value = amax * std::sqrt(ssq);
  } else if (c == 'o' || c == '1') {
    launch_col_norm(m, n, A, lda, 1, 1.0, d_v, s);
    if ((rc = fetch(n))) return rc;
    *value = fold_max(n);
  } else if (c == 'i') {
    launch_row_abssum(m, n, A, lda, d_v, s);
    if ((rc = fetch(m))) return rc;
    *value = fold_max(m);
  }
  return 0;
}

}  // namespace la
