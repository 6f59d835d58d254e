// capi.cpp -- the C ABI of include/lapack_amd.h: Fortran LAPACK drop-ins,
// value-argument entry points and device (async) entry points.
//
// Argument checking follows the reference routine by routine (order and
// parameter numbers): DGETRF SRC/dgetrf.f:144-155, DGETRF2 SRC/dgetrf2.f:152-163,
// DGETRS SRC/dgetrs.f:157-174, DGESV SRC/dgesv.f:148-162, DSGESV
// SRC/dsgesv.f:263-280, DGEMM
// BLAS/SRC/dgemm.f:271-294, DLASWP none (SRC/dlaswp.f:138-150).  The
// Fortran symbols call XERBLA with the routine name and -INFO and return,
// like the reference; the lapack_amd_* entry points return -INFO instead.
#include "la_common.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <vector>

#include "../../include/lapack_amd.h"

namespace la {
struct DevCtx;
template <class T> int getrf_device(int m, int n, T* A, long lda, int* ipiv, int* info, hipStream_t s);
template <class T>
int getrs_device(char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb, hipStream_t s);
template <class T>
int getrs_device_ex(char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb, hipStream_t s,
                    int* waitflag);
int with_safe_forms(const std::function<int()>& f);
void fail_waits(int count);
int device_status(int clear);
int tune_getrf(int n, const char* path, int* nb, int* leaf, double* ms);
int load_tuning(const char* path);
template <class T>
int laswp_device(int ncols, T* A, long lda, int nrows, int k1, int k2, const int* ipiv, int incx, hipStream_t s);
void set_tuning(int nb, int leaf);
int tune_nb();
int tune_leaf();
int tune_nb_for(int n);
int tune_leaf_for(int n);
DevCtx* ctx_for_current();
hipStream_t ctx_stream(DevCtx* c);
std::mutex& io_mutex(DevCtx* c);
int io_buffer(DevCtx* c, size_t bytes, char** out);
bool is_device_ptr(const void* p);
void phase_enable(int on);
void phase_read(double* ms, long* calls, double* flops);
void phase_reset();
int dsgesv_device(int n, int nrhs, double* A, long lda, int* ipiv, const double* B, long ldb, double* X, long ldx,
                  int* iter, hipStream_t s);
int dgesv_ir_device(int n, int nrhs, double* A, long lda, int* ipiv, const double* B, long ldb, double* X, long ldx,
                    int kind, int* iter, int* gmres_iter, double* berr, hipStream_t s);
int geequ_device(int m, int n, const double* A, long lda, double* r, double* c, double* rowcnd, double* colcnd,
                 double* amax, int radix, double* d_r, double* d_c, hipStream_t s);
char laqge_device(int m, int n, double* A, long lda, const double* d_r, const double* d_c, double rowcnd,
                  double colcnd, double amax, hipStream_t s);
int gerfs_device(char trans, int n, int nrhs, const double* A, long lda, const double* AF, long ldaf,
                 const int* ipiv, const double* B, long ldb, double* X, long ldx, double* ferr, double* berr,
                 hipStream_t s);
int gecon_device(bool onenrm, int n, const double* AF, long ldaf, double anorm, double* rcond, hipStream_t s);
int getri_device(int n, double* A, long lda, const int* ipiv, hipStream_t s);
int lange_device(char norm, int m, int n, const double* A, long lda, double* value, hipStream_t s);
}  // namespace la

using namespace la;

namespace {

bool lsame(char a, char b) { return (a | 0x20) == (b | 0x20); }
bool is_trans(char t) { return lsame(t, 'T') || lsame(t, 'C'); }

// ---- XERBLA state ----------------------------------------------------------
std::mutex g_xmu;
bool g_xrecord = false;
char g_xname[16];
int g_xinfo = 0;
int g_xcalls = 0;

void call_xerbla(const char* name, int info) {
  const int p = -info;
  xerbla_(name, &p, std::strlen(name));
}

int hipc(hipError_t e) {
  if (e == hipSuccess) return 0;
  std::fprintf(stderr, "lapack_amd: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
  return LA_ERR_HIP;
}
#define CK(x)                     \
  do {                            \
    int _r = hipc(x);             \
    if (_r) return _r;            \
  } while (0)

size_t al256(size_t x) { return (x + 255) / 256 * 256; }

// One host-or-device array seen through a device pointer.  Host arrays are
// copied into the staging buffer (tightly packed, ld = rows).
template <class T>
struct Arr {
  T* user = nullptr;
  long user_ld = 0;
  int rows = 0, cols = 0;
  T* dev = nullptr;
  long ld = 0;
  bool host = false;
  size_t bytes() const { return host ? al256(sizeof(T) * (size_t)std::max(rows, 1) * std::max(cols, 1)) : 0; }
};
template <class T>
Arr<T> arr(T* p, long ld, int rows, int cols) {
  Arr<T> a;
  a.user = p;
  a.user_ld = ld;
  a.rows = rows;
  a.cols = cols;
  a.host = (p != nullptr) && !is_device_ptr(p);
  a.dev = p;
  a.ld = ld;
  return a;
}
template <class T>
void place(Arr<T>& a, char*& cur) {
  if (!a.host) return;
  a.dev = reinterpret_cast<T*>(cur);
  a.ld = std::max(a.rows, 1);
  cur += a.bytes();
}
template <class T>
int to_dev(Arr<T>& a, hipStream_t s) {
  if (!a.host || a.rows <= 0 || a.cols <= 0) return 0;
  CK(hipMemcpy2DAsync(a.dev, sizeof(T) * a.ld, a.user, sizeof(T) * a.user_ld, sizeof(T) * a.rows, a.cols,
                      hipMemcpyHostToDevice, s));
  return 0;
}
template <class T>
int to_host(Arr<T>& a, hipStream_t s) {
  if (!a.host || a.rows <= 0 || a.cols <= 0) return 0;
  CK(hipMemcpy2DAsync(a.user, sizeof(T) * a.user_ld, a.dev, sizeof(T) * a.ld, sizeof(T) * a.rows, a.cols,
                      hipMemcpyDeviceToHost, s));
  return 0;
}

struct Call {
  DevCtx* c = nullptr;
  hipStream_t s = nullptr;
  std::unique_lock<std::mutex> lk;
  int open() {
    c = ctx_for_current();
    if (!c) return LA_ERR_NODEVICE;
    s = ctx_stream(c);
    lk = std::unique_lock<std::mutex>(io_mutex(c));
    return 0;
  }
};

// ---- generic synchronous drivers (host or device pointers) ----------------

template <class T>
int getrf_sync(int m, int n, T* a, int lda, int* ipiv) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  if (m == 0 || n == 0) return 0;
  Call call;
  int rc = call.open();
  if (rc) return rc;
  const int mn = std::min(m, n);
  Arr<T> A = arr<T>(a, lda, m, n);
  Arr<int> P = arr<int>(ipiv, mn, mn, 1);
  size_t need = A.bytes() + P.bytes() + 256;
  char* io = nullptr;
  rc = io_buffer(call.c, need, &io);
  if (rc) return rc;
  char* cur = io;
  int* dinfo = reinterpret_cast<int*>(cur);
  cur += 256;
  place(A, cur);
  place(P, cur);
  if ((rc = to_dev(A, call.s))) return rc;
  int info = 0;
  auto factor = [&]() -> int {
    int r = getrf_device<T>(m, n, A.dev, A.ld, P.dev, dinfo, call.s);
    if (r) return r;
    if ((r = to_host(A, call.s))) return r;
    if ((r = to_host(P, call.s))) return r;
    CK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, call.s));
    CK(hipStreamSynchronize(call.s));
    return 0;
  };
  if ((rc = factor())) return rc;
  // a bounded wait gave up: A was left as passed; factor it again with the
  // panel form that has no waits
  if (info == LA_ERR_SYNC && (rc = with_safe_forms(factor))) return rc;
  return info;
}

template <class T>
int getrs_check(char trans, int n, int nrhs, int lda, int ldb) {
  if (!lsame(trans, 'N') && !is_trans(trans)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max(1, n)) return -5;
  if (ldb < std::max(1, n)) return -8;
  return 0;
}

template <class T>
int getrs_sync(char trans, int n, int nrhs, const T* a, int lda, const int* ipiv, T* b, int ldb) {
  int rc = getrs_check<T>(trans, n, nrhs, lda, ldb);
  if (rc) return rc;
  if (n == 0 || nrhs == 0) return 0;
  Call call;
  if ((rc = call.open())) return rc;
  Arr<T> A = arr<T>(const_cast<T*>(a), lda, n, n);
  Arr<int> P = arr<int>(const_cast<int*>(ipiv), n, n, 1);
  Arr<T> B = arr<T>(b, ldb, n, nrhs);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + P.bytes() + B.bytes() + 256, &io))) return rc;
  char* cur = io;
  int* dflag = reinterpret_cast<int*>(cur);
  cur += 256;
  place(A, cur);
  place(P, cur);
  place(B, cur);
  if ((rc = to_dev(A, call.s)) || (rc = to_dev(P, call.s)) || (rc = to_dev(B, call.s))) return rc;
  int flag = 0;
  auto solve = [&]() -> int {
    launch_fill_i32(dflag, 1, 0, call.s);
    int r = getrs_device_ex<T>(trans, n, nrhs, A.dev, A.ld, P.dev, B.dev, B.ld, call.s, dflag);
    if (r) return r;
    if ((r = to_host(B, call.s))) return r;
    CK(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, call.s));
    CK(hipStreamSynchronize(call.s));
    return 0;
  };
  if ((rc = solve())) return rc;
  // a bounded wait gave up: B was left as passed; solve again without waits
  if (flag && (rc = with_safe_forms(solve))) return rc;
  return flag ? LA_ERR_SYNC : 0;
}

int gesv_sync(int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb) {
  if (n < 0) return -1;
  if (nrhs < 0) return -2;
  if (lda < std::max(1, n)) return -4;
  if (ldb < std::max(1, n)) return -7;
  if (n == 0) return 0;
  Call call;
  int rc = call.open();
  if (rc) return rc;
  Arr<double> A = arr<double>(a, lda, n, n);
  Arr<int> P = arr<int>(ipiv, n, n, 1);
  Arr<double> B = arr<double>(b, ldb, n, nrhs);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + P.bytes() + B.bytes() + 256, &io))) return rc;
  char* cur = io;
  int* dinfo = reinterpret_cast<int*>(cur);
  cur += 256;
  place(A, cur);
  place(P, cur);
  place(B, cur);
  if ((rc = to_dev(A, call.s)) || (rc = to_dev(B, call.s))) return rc;
  int info = 0, flag = 0;
  auto factor = [&]() -> int {
    int r = getrf_device<double>(n, n, A.dev, A.ld, P.dev, dinfo, call.s);
    if (r) return r;
    CK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, call.s));
    CK(hipStreamSynchronize(call.s));
    return 0;
  };
  auto solve = [&]() -> int {
    launch_fill_i32(dinfo + 1, 1, 0, call.s);
    int r = getrs_device_ex<double>('N', n, nrhs, A.dev, A.ld, P.dev, B.dev, B.ld, call.s, dinfo + 1);
    if (r) return r;
    CK(hipMemcpyAsync(&flag, dinfo + 1, sizeof(int), hipMemcpyDeviceToHost, call.s));
    CK(hipStreamSynchronize(call.s));
    return 0;
  };
  if ((rc = factor())) return rc;
  // a bounded wait gave up: A was left as passed; factor it again without waits
  if (info == LA_ERR_SYNC && (rc = with_safe_forms(factor))) return rc;
  // DGESV solves only when the factorization is nonsingular (SRC/dgesv.f:170-175).
  if (info == 0 && nrhs > 0) {
    if ((rc = solve())) return rc;
    if (flag && (rc = with_safe_forms(solve))) return rc;
    if (flag) info = LA_ERR_SYNC;
  }
  if ((rc = to_host(A, call.s)) || (rc = to_host(P, call.s))) return rc;
  if (info == 0 && (rc = to_host(B, call.s))) return rc;
  CK(hipStreamSynchronize(call.s));
  return info;
}

// DSGESV: arguments checked as SRC/dsgesv.f:263-280.
int dsgesv_sync(int n, int nrhs, double* a, int lda, int* ipiv, const double* b, int ldb, double* x, int ldx,
                int* iter) {
  *iter = 0;
  if (n < 0) return -1;
  if (nrhs < 0) return -2;
  if (lda < std::max(1, n)) return -4;
  if (ldb < std::max(1, n)) return -7;
  if (ldx < std::max(1, n)) return -9;
  if (n == 0) return 0;
  Call call;
  int rc = call.open();
  if (rc) return rc;
  Arr<double> A = arr<double>(a, lda, n, n);
  Arr<int> P = arr<int>(ipiv, n, n, 1);
  Arr<double> B = arr<double>(const_cast<double*>(b), ldb, n, nrhs);
  Arr<double> X = arr<double>(x, ldx, n, nrhs);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + P.bytes() + B.bytes() + X.bytes() + 256, &io))) return rc;
  char* cur = io;
  place(A, cur);
  place(P, cur);
  place(B, cur);
  place(X, cur);
  if ((rc = to_dev(A, call.s)) || (rc = to_dev(B, call.s))) return rc;
  const int info = dsgesv_device(n, nrhs, A.dev, A.ld, P.dev, B.dev, B.ld, X.dev, X.ld, iter, call.s);
  if (info < 0) return info;
  if ((rc = to_host(A, call.s)) || (rc = to_host(P, call.s)) || (rc = to_host(X, call.s))) return rc;
  CK(hipStreamSynchronize(call.s));
  // a bounded wait inside gave up: the results are not to be trusted
  if (device_status(1) > 0) return LA_ERR_SYNC;
  return info;
}

// Mixed-precision solve with refinement / GMRES-IR (no reference routine;
// arguments checked like DSGESV, SRC/dsgesv.f:263-280, PREC as parameter 10).
int dgesv_ir_sync(int n, int nrhs, double* a, int lda, int* ipiv, const double* b, int ldb, double* x, int ldx,
                  char prec, int* iter, int* gmres_iter, double* berr) {
  *iter = 0;
  *gmres_iter = 0;
  if (n < 0) return -1;
  if (nrhs < 0) return -2;
  if (lda < std::max(1, n)) return -4;
  if (ldb < std::max(1, n)) return -7;
  if (ldx < std::max(1, n)) return -9;
  int kind;
  if (lsame(prec, 'S')) kind = 0;
  else if (lsame(prec, 'H')) kind = 1;
  else if (lsame(prec, 'B')) kind = 2;
  else return -10;
  if (n == 0) return 0;
  Call call;
  int rc = call.open();
  if (rc) return rc;
  Arr<double> A = arr<double>(a, lda, n, n);
  Arr<int> P = arr<int>(ipiv, n, n, 1);
  Arr<double> B = arr<double>(const_cast<double*>(b), ldb, n, nrhs);
  Arr<double> X = arr<double>(x, ldx, n, nrhs);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + P.bytes() + B.bytes() + X.bytes() + 256, &io))) return rc;
  char* cur = io;
  place(A, cur);
  place(P, cur);
  place(B, cur);
  place(X, cur);
  if ((rc = to_dev(A, call.s)) || (rc = to_dev(B, call.s))) return rc;
  const int info = dgesv_ir_device(n, nrhs, A.dev, A.ld, P.dev, B.dev, B.ld, X.dev, X.ld, kind, iter, gmres_iter,
                                   berr, call.s);
  if (info < 0) return info;
  if ((rc = to_host(A, call.s)) || (rc = to_host(P, call.s)) || (rc = to_host(X, call.s))) return rc;
  CK(hipStreamSynchronize(call.s));
  if (device_status(1) > 0) return LA_ERR_SYNC;
  return info;
}

// DGERFS (argument checks SRC/dgerfs.f:236-256).
int gerfs_sync(char trans, int n, int nrhs, const double* a, int lda, const double* af, int ldaf, const int* ipiv,
               const double* b, int ldb, double* x, int ldx, double* ferr, double* berr) {
  if (!lsame(trans, 'N') && !is_trans(trans)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max(1, n)) return -5;
  if (ldaf < std::max(1, n)) return -7;
  if (ldb < std::max(1, n)) return -10;
  if (ldx < std::max(1, n)) return -12;
  std::vector<double> hf(std::max(nrhs, 1)), hb(std::max(nrhs, 1));
  const bool fdev = is_device_ptr(ferr), bdev = is_device_ptr(berr);
  if (n == 0 || nrhs == 0) {
    for (int j = 0; j < nrhs; ++j) { hf[j] = 0.0; hb[j] = 0.0; }
  } else {
    Call call;
    int rc = call.open();
    if (rc) return rc;
    Arr<double> A = arr<double>(const_cast<double*>(a), lda, n, n);
    Arr<double> F = arr<double>(const_cast<double*>(af), ldaf, n, n);
    Arr<int> P = arr<int>(const_cast<int*>(ipiv), n, n, 1);
    Arr<double> B = arr<double>(const_cast<double*>(b), ldb, n, nrhs);
    Arr<double> X = arr<double>(x, ldx, n, nrhs);
    char* io = nullptr;
    if ((rc = io_buffer(call.c, A.bytes() + F.bytes() + P.bytes() + B.bytes() + X.bytes() + 256, &io))) return rc;
    char* cur = io;
    place(A, cur);
    place(F, cur);
    place(P, cur);
    place(B, cur);
    place(X, cur);
    if ((rc = to_dev(A, call.s)) || (rc = to_dev(F, call.s)) || (rc = to_dev(P, call.s)) ||
        (rc = to_dev(B, call.s)) || (rc = to_dev(X, call.s)))
      return rc;
    const char tr = lsame(trans, 'N') ? 'N' : 'T';
    if ((rc = gerfs_device(tr, n, nrhs, A.dev, A.ld, F.dev, F.ld, P.dev, B.dev, B.ld, X.dev, X.ld, hf.data(),
                           hb.data(), call.s)))
      return rc;
    if ((rc = to_host(X, call.s))) return rc;
    CK(hipStreamSynchronize(call.s));
    if (device_status(1) > 0) return LA_ERR_SYNC;
  }
  if (nrhs > 0) {
    if (fdev) CK(hipMemcpy(ferr, hf.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice));
    else std::memcpy(ferr, hf.data(), sizeof(double) * nrhs);
    if (bdev) CK(hipMemcpy(berr, hb.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice));
    else std::memcpy(berr, hb.data(), sizeof(double) * nrhs);
  }
  return 0;
}

// A host-or-device array of `count` doubles read into / written from host memory.
int get_hd(const double* p, double* out, int count) {
  if (count <= 0) return 0;
  if (is_device_ptr(p)) CK(hipMemcpy(out, p, sizeof(double) * count, hipMemcpyDeviceToHost));
  else std::memcpy(out, p, sizeof(double) * count);
  return 0;
}
int put_hd(double* p, const double* in, int count) {
  if (count <= 0) return 0;
  if (is_device_ptr(p)) CK(hipMemcpy(p, in, sizeof(double) * count, hipMemcpyHostToDevice));
  else std::memcpy(p, in, sizeof(double) * count);
  return 0;
}

// DLANGE (SRC/dlange.f:114-200): returns the norm; *err receives a library
// error code (the reference has no error exit).
double lange_sync(char norm, int m, int n, const double* a, int lda, int* err) {
  *err = 0;
  if (std::min(m, n) <= 0) return 0.0;
  Call call;
  int rc = call.open();
  if (rc) { *err = rc; return 0.0; }
  Arr<double> A = arr<double>(const_cast<double*>(a), lda, m, n);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + 256, &io))) { *err = rc; return 0.0; }
  char* cur = io;
  place(A, cur);
  if ((rc = to_dev(A, call.s))) { *err = rc; return 0.0; }
  double v = 0.0;
  if ((rc = lange_device(norm, m, n, A.dev, A.ld, &v, call.s))) *err = rc;
  return v;
}

// DGECON (argument checks SRC/dgecon.f:157-185).
int gecon_sync(char norm, int n, const double* a, int lda, double anorm, double* rcond) {
  const bool onenrm = norm == '1' || lsame(norm, 'O');
  if (!onenrm && !lsame(norm, 'I')) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, n)) return -4;
  if (anorm < 0.0) return -5;
  *rcond = 0.0;
  if (n == 0) {
    *rcond = 1.0;
    return 0;
  }
  if (anorm == 0.0) return 0;
  if (std::isnan(anorm)) {
    *rcond = anorm;
    return -5;
  }
  if (anorm > 1.79769313486231571e308) return -5;
  Call call;
  int rc = call.open();
  if (rc) return rc;
  Arr<double> A = arr<double>(const_cast<double*>(a), lda, n, n);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + 256, &io))) return rc;
  char* cur = io;
  place(A, cur);
  if ((rc = to_dev(A, call.s))) return rc;
  return gecon_device(onenrm, n, A.dev, A.ld, anorm, rcond, call.s);
}

// DGETRI (argument checks SRC/dgetri.f:141-158); WORK(1) = N*NB (NB = 64,
// the reference's ILAENV value) as the reference reports its optimal size.
int getri_sync(int n, double* a, int lda, const int* ipiv, double* work, int lwork) {
  const double lwkopt = (double)std::max(1, n * 64);
  if (work) put_hd(work, &lwkopt, 1);
  const bool lquery = lwork == -1;
  if (n < 0) return -1;
  if (lda < std::max(1, n)) return -3;
  if (lwork < std::max(1, n) && !lquery) return -6;
  if (lquery || n == 0) return 0;
  Call call;
  int rc = call.open();
  if (rc) return rc;
  Arr<double> A = arr<double>(a, lda, n, n);
  Arr<int> P = arr<int>(const_cast<int*>(ipiv), n, n, 1);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + P.bytes() + 256, &io))) return rc;
  char* cur = io;
  place(A, cur);
  place(P, cur);
  if ((rc = to_dev(A, call.s)) || (rc = to_dev(P, call.s))) return rc;
  const int info = getri_device(n, A.dev, A.ld, P.dev, call.s);
  if (info < 0) return info;
  if (info == 0 && (rc = to_host(A, call.s))) return rc;
  CK(hipStreamSynchronize(call.s));
  return info;
}

// DGESVX (SRC/dgesvx.f:344-600): the reference's sequence -- equilibration
// (DGEEQU / DLAQGE), DGETRF on a copy, the reciprocal pivot growth, DGECON,
// DGETRS, DGERFS, the back-scaling of X and FERR, INFO = N+1 below eps --
// with every O(n^2) and O(n^3) step on the GPU and the scalar decisions on
// the host.  Matrices may be host or device memory; R, C, FERR, BERR, WORK
// and RCOND too.  Returns INFO; *equed as the reference sets it.
int gesvx_sync(char fact, char trans, int n, int nrhs, double* a, int lda, double* af, int ldaf, int* ipiv,
               char* equed, double* r, double* c, double* b, int ldb, double* x, int ldx, double* rcond,
               double* ferr, double* berr, double* work) {
  const bool nofact = lsame(fact, 'N'), equil = lsame(fact, 'E'), notran = lsame(trans, 'N');
  const double smlnum = 0x1p-1022, bignum = 1.0 / smlnum, eps = 0x1p-53;
  bool rowequ = false, colequ = false;
  double rowcnd = 1.0, colcnd = 1.0, amax = 0.0;
  std::vector<double> hr(std::max(n, 1)), hc(std::max(n, 1));
  if (nofact || equil) {
    *equed = 'N';
  } else {
    rowequ = lsame(*equed, 'R') || lsame(*equed, 'B');
    colequ = lsame(*equed, 'C') || lsame(*equed, 'B');
  }
  if (!nofact && !equil && !lsame(fact, 'F')) return -1;
  if (!notran && !lsame(trans, 'T') && !lsame(trans, 'C')) return -2;
  if (n < 0) return -3;
  if (nrhs < 0) return -4;
  if (lda < std::max(1, n)) return -6;
  if (ldaf < std::max(1, n)) return -8;
  if (lsame(fact, 'F') && !(rowequ || colequ || lsame(*equed, 'N'))) return -10;
  int rc;
  if (rowequ) {
    if ((rc = get_hd(r, hr.data(), n))) return rc;
    double rcmin = bignum, rcmax = 0.0;
    for (int j = 0; j < n; ++j) { rcmin = std::min(rcmin, hr[j]); rcmax = std::max(rcmax, hr[j]); }
    if (rcmin <= 0.0) return -11;
    rowcnd = n > 0 ? std::max(rcmin, smlnum) / std::min(rcmax, bignum) : 1.0;
  }
  if (colequ) {
    if ((rc = get_hd(c, hc.data(), n))) return rc;
    double rcmin = bignum, rcmax = 0.0;
    for (int j = 0; j < n; ++j) { rcmin = std::min(rcmin, hc[j]); rcmax = std::max(rcmax, hc[j]); }
    if (rcmin <= 0.0) return -12;
    colcnd = n > 0 ? std::max(rcmin, smlnum) / std::min(rcmax, bignum) : 1.0;
  }
  if (ldb < std::max(1, n)) return -14;
  if (ldx < std::max(1, n)) return -16;
  std::vector<double> hf(std::max(nrhs, 1), 0.0), hb(std::max(nrhs, 1), 0.0);
  if (n == 0) {
    // quick return as the reference's zero-order path: RCOND from DGECON = 1,
    // FERR = BERR = 0 from DGERFS, RPVGRW = 1
    double one = 1.0;
    put_hd(rcond, &one, 1);
    put_hd(ferr, hf.data(), nrhs);
    put_hd(berr, hb.data(), nrhs);
    if (work) put_hd(work, &one, 1);
    return 0;
  }
  Call call;
  if ((rc = call.open())) return rc;
  Arr<double> A = arr<double>(a, lda, n, n);
  Arr<double> F = arr<double>(af, ldaf, n, n);
  Arr<int> P = arr<int>(ipiv, n, n, 1);
  Arr<double> B = arr<double>(b, ldb, n, nrhs);
  Arr<double> X = arr<double>(x, ldx, n, nrhs);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + F.bytes() + P.bytes() + B.bytes() + X.bytes() +
                                  2 * al256(sizeof(double) * (size_t)n) + 256, &io)))
    return rc;
  char* cur = io;
  int* dinfo = reinterpret_cast<int*>(cur);
  cur += 256;
  double* d_r = reinterpret_cast<double*>(cur);
  cur += al256(sizeof(double) * (size_t)n);
  double* d_c = reinterpret_cast<double*>(cur);
  cur += al256(sizeof(double) * (size_t)n);
  place(A, cur);
  place(F, cur);
  place(P, cur);
  place(B, cur);
  place(X, cur);
  hipStream_t s = call.s;
  if ((rc = to_dev(A, s)) || (rc = to_dev(B, s))) return rc;
  if (!nofact && !equil && ((rc = to_dev(F, s)) || (rc = to_dev(P, s)))) return rc;
  bool a_changed = false, b_changed = false;
  if (equil) {
    const int infequ = geequ_device(n, n, A.dev, A.ld, hr.data(), hc.data(), &rowcnd, &colcnd, &amax, 0, d_r, d_c,
                                    s);
    if (infequ < 0) return infequ;
    // DGEEQU leaves R (and C) as the reference leaves them
    if ((rc = put_hd(r, hr.data(), n))) return rc;
    if ((infequ == 0 || infequ > n) && (rc = put_hd(c, hc.data(), n))) return rc;
    if (infequ == 0) {
      *equed = laqge_device(n, n, A.dev, A.ld, d_r, d_c, rowcnd, colcnd, amax, s);
      a_changed = *equed != 'N';
      rowequ = lsame(*equed, 'R') || lsame(*equed, 'B');
      colequ = lsame(*equed, 'C') || lsame(*equed, 'B');
    }
  } else {
    if (rowequ) CK(hipMemcpyAsync(d_r, hr.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (colequ) CK(hipMemcpyAsync(d_c, hc.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  }
  // B <- diag(R) B or diag(C) B (dgesvx.f:411-425)
  if (notran && rowequ) { launch_scale_rc(n, nrhs, B.dev, B.ld, d_r, nullptr, 1, s); b_changed = true; }
  if (!notran && colequ) { launch_scale_rc(n, nrhs, B.dev, B.ld, d_c, nullptr, 1, s); b_changed = true; }
  auto copy_back = [&](bool with_af, bool with_x) -> int {
    int r2;
    if (a_changed && (r2 = to_host(A, s))) return r2;
    if (b_changed && (r2 = to_host(B, s))) return r2;
    if (with_af && ((r2 = to_host(F, s)) || (r2 = to_host(P, s)))) return r2;
    if (with_x && (r2 = to_host(X, s))) return r2;
    CK(hipStreamSynchronize(s));
    return 0;
  };
  auto upper_max = [&](const double* M, long ld, int rows, int cols, int mode, double* out) -> int {
    double* d_v = d_r;  // n doubles of scratch once R has been used
    std::vector<double> h(cols);
    launch_col_norm(rows, cols, M, ld, mode, 1.0, d_v, s);
    CK(hipMemcpyAsync(h.data(), d_v, sizeof(double) * cols, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    double v = 0.0;
    for (int j = 0; j < cols; ++j)
      if (v < h[j] || std::isnan(h[j])) v = h[j];
    *out = v;
    return 0;
  };
  int info = 0;
  if (nofact || equil) {
    launch_copy_d(n, n, A.dev, A.ld, F.dev, F.ld, s);
    auto factor = [&]() -> int {
      int r2 = getrf_device<double>(n, n, F.dev, F.ld, P.dev, dinfo, s);
      if (r2) return r2;
      CK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
      CK(hipStreamSynchronize(s));
      return 0;
    };
    if ((rc = factor())) return rc;
    if (info == LA_ERR_SYNC && (rc = with_safe_forms(factor))) return rc;
    if (info < 0) return info;
    if (info > 0) {
      // the reciprocal pivot growth of the leading INFO columns (dgesvx.f:437-450)
      double rpvgrw = 0.0, am = 0.0;
      if ((rc = upper_max(F.dev, F.ld, info, info, 3, &rpvgrw))) return rc;
      if (rpvgrw == 0.0) {
        rpvgrw = 1.0;
      } else {
        if ((rc = upper_max(A.dev, A.ld, n, info, 0, &am))) return rc;
        rpvgrw = am / rpvgrw;
      }
      if (work && (rc = put_hd(work, &rpvgrw, 1))) return rc;
      const double zero = 0.0;
      if ((rc = put_hd(rcond, &zero, 1))) return rc;
      if ((rc = copy_back(true, false))) return rc;
      return info;
    }
  }
  double anorm = 0.0, rpvgrw = 0.0, amx = 0.0;
  if ((rc = lange_device(notran ? '1' : 'I', n, n, A.dev, A.ld, &anorm, s))) return rc;
  if ((rc = upper_max(F.dev, F.ld, n, n, 3, &rpvgrw))) return rc;
  if (rpvgrw == 0.0) {
    rpvgrw = 1.0;
  } else {
    if ((rc = upper_max(A.dev, A.ld, n, n, 0, &amx))) return rc;
    rpvgrw = amx / rpvgrw;
  }
  double rc_est = 0.0;
  rc = gecon_device(notran, n, F.dev, F.ld, anorm, &rc_est, s);
  if (rc < 0) return rc;
  launch_copy_d(n, nrhs, B.dev, B.ld, X.dev, X.ld, s);
  if ((rc = getrs_device<double>(notran ? 'N' : 'T', n, nrhs, F.dev, F.ld, P.dev, X.dev, X.ld, s))) return rc;
  if ((rc = gerfs_device(notran ? 'N' : 'T', n, nrhs, A.dev, A.ld, F.dev, F.ld, P.dev, B.dev, B.ld, X.dev, X.ld,
                         hf.data(), hb.data(), s)))
    return rc;
  // X <- diag(C) X, FERR / COLCND (or R, ROWCND for the transposed system)
  if (notran && colequ) {
    launch_scale_rc(n, nrhs, X.dev, X.ld, d_c, nullptr, 1, s);
    for (int j = 0; j < nrhs; ++j) hf[j] = hf[j] / colcnd;
  } else if (!notran && rowequ) {
    CK(hipMemcpyAsync(d_r, hr.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    launch_scale_rc(n, nrhs, X.dev, X.ld, d_r, nullptr, 1, s);
    for (int j = 0; j < nrhs; ++j) hf[j] = hf[j] / rowcnd;
  }
  if ((rc = copy_back(nofact || equil, true))) return rc;
  if (device_status(1) > 0) return LA_ERR_SYNC;
  if ((rc = put_hd(rcond, &rc_est, 1)) || (rc = put_hd(ferr, hf.data(), nrhs)) ||
      (rc = put_hd(berr, hb.data(), nrhs)))
    return rc;
  if (work && (rc = put_hd(work, &rpvgrw, 1))) return rc;
  if (rc_est < eps) info = n + 1;
  return info;
}

// DGEEQU / DGEEQUB (SRC/dgeequ.f:163-173, SRC/dgeequb.f:170-180 argument checks).
int geequ_sync(int m, int n, const double* a, int lda, double* r, double* c, double* rowcnd, double* colcnd,
               double* amax, int radix) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  if (m == 0 || n == 0) {
    *rowcnd = 1.0;
    *colcnd = 1.0;
    *amax = 0.0;
    return 0;
  }
  Call call;
  int rc = call.open();
  if (rc) return rc;
  Arr<double> A = arr<double>(const_cast<double*>(a), lda, m, n);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + al256(sizeof(double) * (size_t)(m + n)) + 256, &io))) return rc;
  char* cur = io;
  double* d_r = reinterpret_cast<double*>(cur);
  double* d_c = d_r + m;
  cur += al256(sizeof(double) * (size_t)(m + n));
  place(A, cur);
  if ((rc = to_dev(A, call.s))) return rc;
  const bool rdev = is_device_ptr(r), cdev = is_device_ptr(c);
  std::vector<double> hr(m), hc(n);
  const int info = geequ_device(m, n, A.dev, A.ld, hr.data(), hc.data(), rowcnd, colcnd, amax, radix, d_r, d_c,
                                call.s);
  if (info < 0) return info;
  // R and C are left as the reference leaves them: on a zero row (INFO <= M)
  // R holds the row maxima and C is untouched; on a zero column R is final
  // and C holds the column maxima.
  if (rdev) CK(hipMemcpy(r, hr.data(), sizeof(double) * m, hipMemcpyHostToDevice));
  else std::memcpy(r, hr.data(), sizeof(double) * m);
  if (info == 0 || info > m) {
    if (cdev) CK(hipMemcpy(c, hc.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    else std::memcpy(c, hc.data(), sizeof(double) * n);
  }
  return info;
}

char laqge_sync(int m, int n, double* a, int lda, const double* r, const double* c, double rowcnd, double colcnd,
                double amax, int* err) {
  *err = 0;
  if (m <= 0 || n <= 0) return 'N';
  Call call;
  int rc = call.open();
  if (rc) { *err = rc; return 'N'; }
  Arr<double> A = arr<double>(a, lda, m, n);
  Arr<double> R = arr<double>(const_cast<double*>(r), m, m, 1);
  Arr<double> C = arr<double>(const_cast<double*>(c), n, n, 1);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + R.bytes() + C.bytes() + 256, &io))) { *err = rc; return 'N'; }
  char* cur = io;
  place(A, cur);
  place(R, cur);
  place(C, cur);
  if ((rc = to_dev(A, call.s)) || (rc = to_dev(R, call.s)) || (rc = to_dev(C, call.s))) { *err = rc; return 'N'; }
  const char equed = laqge_device(m, n, A.dev, A.ld, R.dev, C.dev, rowcnd, colcnd, amax, call.s);
  if (equed != 'N' && (rc = to_host(A, call.s))) { *err = rc; return equed; }
  if ((rc = hipc(hipStreamSynchronize(call.s)))) *err = rc;
  return equed;
}

template <class T>
int laswp_sync(int n, T* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  if (n <= 0 || incx == 0) return 0;
  Call call;
  int rc = call.open();
  if (rc) return rc;
  // Gather the IPIV entries the reference reads (IX0 + (I-I1)*INCX, SRC/dlaswp.f:138-150)
  // into a contiguous host vector, and find how far down the rows go.
  const bool dip = is_device_ptr(ipiv);
  const int cnt = k2 - k1 + 1;
  if (cnt <= 0) return 0;
  const int ainc = incx > 0 ? incx : -incx;
  const long span = (long)(cnt - 1) * ainc + 1;
  std::vector<int> raw(span), piv(cnt);
  if (dip) CK(hipMemcpy(raw.data(), ipiv + (k1 - 1), sizeof(int) * span, hipMemcpyDeviceToHost));
  else std::memcpy(raw.data(), ipiv + (k1 - 1), sizeof(int) * span);
  int maxrow = k2;
  for (int i = 0; i < cnt; ++i) {
    piv[i] = raw[(long)i * ainc];
    maxrow = std::max(maxrow, piv[i]);
  }
  // In both directions the reference pairs row I with IPIV(K1+(I-K1)*|INCX|)
  // (IX0 = K1+(K1-K2)*INCX for INCX < 0); only the order of the walk differs.
  const int nrows = std::min(maxrow, lda);
  Arr<T> A = arr<T>(a, lda, nrows, n);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + al256(sizeof(int) * (size_t)k2) + 256, &io))) return rc;
  char* cur = io;
  int* dpiv = reinterpret_cast<int*>(cur);
  cur += al256(sizeof(int) * (size_t)k2);
  place(A, cur);
  // device IPIV laid out so that dpiv[k-1] is the pivot paired with row k
  std::vector<int> full(k2, 0);
  for (int i = 0; i < cnt; ++i) full[k1 - 1 + i] = piv[i];
  CK(hipMemcpyAsync(dpiv, full.data(), sizeof(int) * (size_t)k2, hipMemcpyHostToDevice, call.s));
  if ((rc = to_dev(A, call.s))) return rc;
  if ((rc = laswp_device<T>(n, A.dev, A.ld, nrows, k1, k2, dpiv, incx > 0 ? 1 : -1, call.s))) return rc;
  if ((rc = to_host(A, call.s))) return rc;
  CK(hipStreamSynchronize(call.s));
  return 0;
}

int gemm_check(char ta, char tb, int m, int n, int k, int lda, int ldb, int ldc) {
  const bool nota = lsame(ta, 'N'), notb = lsame(tb, 'N');
  const int nrowa = nota ? m : k, nrowb = notb ? k : n;
  if (!nota && !is_trans(ta)) return 1;
  if (!notb && !is_trans(tb)) return 2;
  if (m < 0) return 3;
  if (n < 0) return 4;
  if (k < 0) return 5;
  if (lda < std::max(1, nrowa)) return 8;
  if (ldb < std::max(1, nrowb)) return 10;
  if (ldc < std::max(1, m)) return 13;
  return 0;
}

int gemm_sync(char ta, char tb, int m, int n, int k, double alpha, const double* a, int lda, const double* b,
              int ldb, double beta, double* c, int ldc) {
  int e = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
  if (e) return -e;
  if (m == 0 || n == 0) return 0;
  Call call;
  int rc = call.open();
  if (rc) return rc;
  const bool nota = lsame(ta, 'N'), notb = lsame(tb, 'N');
  Arr<double> A = arr<double>(const_cast<double*>(a), lda, nota ? m : k, nota ? k : m);
  Arr<double> B = arr<double>(const_cast<double*>(b), ldb, notb ? k : n, notb ? n : k);
  Arr<double> C = arr<double>(c, ldc, m, n);
  char* io = nullptr;
  if ((rc = io_buffer(call.c, A.bytes() + B.bytes() + C.bytes() + 256, &io))) return rc;
  char* cur = io;
  place(A, cur);
  place(B, cur);
  place(C, cur);
  if ((rc = to_dev(A, call.s)) || (rc = to_dev(B, call.s)) || (rc = to_dev(C, call.s))) return rc;
  launch_dgemm(ta, tb, m, n, k, alpha, A.dev, A.ld, B.dev, B.ld, beta, C.dev, C.ld, call.s);
  CK(hipGetLastError());
  if ((rc = to_host(C, call.s))) return rc;
  CK(hipStreamSynchronize(call.s));
  return 0;
}

}  // namespace

namespace la {
// DGETRS's argument checks (SRC/dgetrs.f:157-174) for other sources.
int getrs_arg_check(char trans, int n, int nrhs, int lda, int ldb) {
  return getrs_check<double>(trans, n, nrhs, lda, ldb);
}
}  // namespace la

extern "C" {

// ---- XERBLA -----------------------------------------------------------------
// Reference behaviour (SRC/xerbla.f:86-88): print, then STOP.  Weak, so a
// program's own XERBLA takes precedence at link time.
__attribute__((weak)) void xerbla_(const char* srname, const int* info, size_t len) {
  int n = (int)len;
  while (n > 0 && srname[n - 1] == ' ') --n;
  {
    std::lock_guard<std::mutex> g(g_xmu);
    const int k = std::min(n, (int)sizeof(g_xname) - 1);
    std::memcpy(g_xname, srname, k);
    g_xname[k] = 0;
    g_xinfo = *info;
    g_xcalls++;
    if (g_xrecord) return;
  }
  std::printf(" ** On entry to %.*s parameter number %2d had an illegal value\n", n, srname, *info);
  std::fflush(stdout);
  std::exit(0);
}

void lapack_amd_xerbla_record(int on) {
  std::lock_guard<std::mutex> g(g_xmu);
  g_xrecord = on != 0;
  g_xcalls = 0;
  g_xinfo = 0;
  g_xname[0] = 0;
}
int lapack_amd_xerbla_last(char* name_out, int cap) {
  std::lock_guard<std::mutex> g(g_xmu);
  if (name_out && cap > 0) {
    std::strncpy(name_out, g_xname, (size_t)cap - 1);
    name_out[cap - 1] = 0;
  }
  return g_xinfo;
}
int lapack_amd_xerbla_calls(void) {
  std::lock_guard<std::mutex> g(g_xmu);
  return g_xcalls;
}

// ---- Fortran drop-ins -----------------------------------------------------------
void dgetrf_(const int* m, const int* n, double* a, const int* lda, int* ipiv, int* info) {
  *info = getrf_sync<double>(*m, *n, a, *lda, ipiv);
  if (*info < 0 && *info > -1000) call_xerbla("DGETRF", *info);
}
void dgetrf2_(const int* m, const int* n, double* a, const int* lda, int* ipiv, int* info) {
  *info = getrf_sync<double>(*m, *n, a, *lda, ipiv);
  if (*info < 0 && *info > -1000) call_xerbla("DGETRF2", *info);
}
void sgetrf_(const int* m, const int* n, float* a, const int* lda, int* ipiv, int* info) {
  *info = getrf_sync<float>(*m, *n, a, *lda, ipiv);
  if (*info < 0 && *info > -1000) call_xerbla("SGETRF", *info);
}
void dgetrs_(const char* trans, const int* n, const int* nrhs, const double* a, const int* lda, const int* ipiv,
             double* b, const int* ldb, int* info, size_t) {
  *info = getrs_sync<double>(*trans, *n, *nrhs, a, *lda, ipiv, b, *ldb);
  if (*info < 0 && *info > -1000) call_xerbla("DGETRS", *info);
}
void sgetrs_(const char* trans, const int* n, const int* nrhs, const float* a, const int* lda, const int* ipiv,
             float* b, const int* ldb, int* info, size_t) {
  *info = getrs_sync<float>(*trans, *n, *nrhs, a, *lda, ipiv, b, *ldb);
  if (*info < 0 && *info > -1000) call_xerbla("SGETRS", *info);
}
void dgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b, const int* ldb,
            int* info) {
  *info = gesv_sync(*n, *nrhs, a, *lda, ipiv, b, *ldb);
  if (*info < 0 && *info > -1000) call_xerbla("DGESV ", *info);
}
void dsgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, const double* b,
             const int* ldb, double* x, const int* ldx, double*, float*, int* iter, int* info) {
  *info = dsgesv_sync(*n, *nrhs, a, *lda, ipiv, b, *ldb, x, *ldx, iter);
  if (*info < 0 && *info > -1000) call_xerbla("DSGESV", *info);
}
void dgerfs_(const char* trans, const int* n, const int* nrhs, const double* a, const int* lda, const double* af,
             const int* ldaf, const int* ipiv, const double* b, const int* ldb, double* x, const int* ldx,
             double* ferr, double* berr, double*, int*, int* info, size_t) {
  *info = gerfs_sync(*trans, *n, *nrhs, a, *lda, af, *ldaf, ipiv, b, *ldb, x, *ldx, ferr, berr);
  if (*info < 0 && *info > -1000) call_xerbla("DGERFS", *info);
}
void dgeequ_(const int* m, const int* n, const double* a, const int* lda, double* r, double* c, double* rowcnd,
             double* colcnd, double* amax, int* info) {
  *info = geequ_sync(*m, *n, a, *lda, r, c, rowcnd, colcnd, amax, 0);
  if (*info < 0 && *info > -1000) call_xerbla("DGEEQU", *info);
}
void dgeequb_(const int* m, const int* n, const double* a, const int* lda, double* r, double* c, double* rowcnd,
              double* colcnd, double* amax, int* info) {
  *info = geequ_sync(*m, *n, a, *lda, r, c, rowcnd, colcnd, amax, 1);
  if (*info < 0 && *info > -1000) call_xerbla("DGEEQUB", *info);
}
void dlaqge_(const int* m, const int* n, double* a, const int* lda, const double* r, const double* c,
             const double* rowcnd, const double* colcnd, const double* amax, char* equed, size_t) {
  int err = 0;
  *equed = laqge_sync(*m, *n, a, *lda, r, c, *rowcnd, *colcnd, *amax, &err);
}
double dlange_(const char* norm, const int* m, const int* n, const double* a, const int* lda, double*, size_t) {
  int err = 0;
  return lange_sync(*norm, *m, *n, a, *lda, &err);
}
void dgecon_(const char* norm, const int* n, const double* a, const int* lda, const double* anorm, double* rcond,
             double*, int*, int* info, size_t) {
  *info = gecon_sync(*norm, *n, a, *lda, *anorm, rcond);
  if (*info < 0 && *info > -1000) call_xerbla("DGECON", *info);
}
void dgetri_(const int* n, double* a, const int* lda, const int* ipiv, double* work, const int* lwork, int* info) {
  *info = getri_sync(*n, a, *lda, ipiv, work, *lwork);
  if (*info < 0 && *info > -1000) call_xerbla("DGETRI", *info);
}
void dgesvx_(const char* fact, const char* trans, const int* n, const int* nrhs, double* a, const int* lda, double* af,
             const int* ldaf, int* ipiv, char* equed, double* r, double* c, double* b, const int* ldb, double* x,
             const int* ldx, double* rcond, double* ferr, double* berr, double* work, int*, int* info, size_t, size_t,
             size_t) {
  *info = gesvx_sync(*fact, *trans, *n, *nrhs, a, *lda, af, *ldaf, ipiv, equed, r, c, b, *ldb, x, *ldx, rcond, ferr,
                     berr, work);
  if (*info < 0 && *info > -1000) call_xerbla("DGESVX", *info);
}
void dlaswp_(const int* n, double* a, const int* lda, const int* k1, const int* k2, const int* ipiv,
             const int* incx) {
  (void)laswp_sync<double>(*n, a, *lda, *k1, *k2, ipiv, *incx);
}
void dgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const double* alpha,
            const double* a, const int* lda, const double* b, const int* ldb, const double* beta, double* c,
            const int* ldc, size_t, size_t) {
  const int rc = gemm_sync(*transa, *transb, *m, *n, *k, *alpha, a, *lda, b, *ldb, *beta, c, *ldc);
  if (rc < 0 && rc > -1000) call_xerbla("DGEMM ", rc);
}

// ---- value-argument entry points ------------------------------------------
int lapack_amd_dgetrf(int m, int n, double* a, int lda, int* ipiv) { return getrf_sync<double>(m, n, a, lda, ipiv); }
int lapack_amd_sgetrf(int m, int n, float* a, int lda, int* ipiv) { return getrf_sync<float>(m, n, a, lda, ipiv); }
int lapack_amd_dgetrs(char trans, int n, int nrhs, const double* a, int lda, const int* ipiv, double* b, int ldb) {
  return getrs_sync<double>(trans, n, nrhs, a, lda, ipiv, b, ldb);
}
int lapack_amd_sgetrs(char trans, int n, int nrhs, const float* a, int lda, const int* ipiv, float* b, int ldb) {
  return getrs_sync<float>(trans, n, nrhs, a, lda, ipiv, b, ldb);
}
int lapack_amd_dgesv(int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb) {
  return gesv_sync(n, nrhs, a, lda, ipiv, b, ldb);
}
int lapack_amd_dsgesv(int n, int nrhs, double* a, int lda, int* ipiv, const double* b, int ldb, double* x,
                      int ldx, int* iter) {
  return dsgesv_sync(n, nrhs, a, lda, ipiv, b, ldb, x, ldx, iter);
}
int lapack_amd_dgesv_ir(int n, int nrhs, double* a, int lda, int* ipiv, const double* b, int ldb, double* x, int ldx,
                        char prec, int* iter, int* gmres_iter, double* berr) {
  return dgesv_ir_sync(n, nrhs, a, lda, ipiv, b, ldb, x, ldx, prec, iter, gmres_iter, berr);
}
int lapack_amd_dgerfs(char trans, int n, int nrhs, const double* a, int lda, const double* af, int ldaf,
                      const int* ipiv, const double* b, int ldb, double* x, int ldx, double* ferr, double* berr) {
  return gerfs_sync(trans, n, nrhs, a, lda, af, ldaf, ipiv, b, ldb, x, ldx, ferr, berr);
}
int lapack_amd_dgeequ(int m, int n, const double* a, int lda, double* r, double* c, double* rowcnd, double* colcnd,
                      double* amax) {
  return geequ_sync(m, n, a, lda, r, c, rowcnd, colcnd, amax, 0);
}
int lapack_amd_dgeequb(int m, int n, const double* a, int lda, double* r, double* c, double* rowcnd, double* colcnd,
                       double* amax) {
  return geequ_sync(m, n, a, lda, r, c, rowcnd, colcnd, amax, 1);
}
char lapack_amd_dlaqge(int m, int n, double* a, int lda, const double* r, const double* c, double rowcnd,
                       double colcnd, double amax) {
  int err = 0;
  return laqge_sync(m, n, a, lda, r, c, rowcnd, colcnd, amax, &err);
}
double lapack_amd_dlange(char norm, int m, int n, const double* a, int lda) {
  int err = 0;
  const double v = lange_sync(norm, m, n, a, lda, &err);
  return err ? std::nan("") : v;
}
int lapack_amd_dgecon(char norm, int n, const double* a, int lda, double anorm, double* rcond) {
  return gecon_sync(norm, n, a, lda, anorm, rcond);
}
int lapack_amd_dgetri(int n, double* a, int lda, const int* ipiv) {
  double w = 0.0;
  return getri_sync(n, a, lda, ipiv, &w, std::max(1, n));
}
int lapack_amd_dgesvx(char fact, char trans, int n, int nrhs, double* a, int lda, double* af, int ldaf, int* ipiv,
                      char* equed, double* r, double* c, double* b, int ldb, double* x, int ldx, double* rcond,
                      double* ferr, double* berr, double* rpvgrw) {
  return gesvx_sync(fact, trans, n, nrhs, a, lda, af, ldaf, ipiv, equed, r, c, b, ldb, x, ldx, rcond, ferr, berr,
                    rpvgrw);
}
int lapack_amd_dlaswp(int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  return laswp_sync<double>(n, a, lda, k1, k2, ipiv, incx);
}
int lapack_amd_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* a, int lda, const double* b,
                     int ldb, double beta, double* c, int ldc) {
  return gemm_sync(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

// ---- device entry points ----------------------------------------------------
int lapack_amd_dgetrf_async(int m, int n, double* d_a, int lda, int* d_ipiv, int* d_info, void* stream) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  return getrf_device<double>(m, n, d_a, lda, d_ipiv, d_info, (hipStream_t)stream);
}
int lapack_amd_sgetrf_async(int m, int n, float* d_a, int lda, int* d_ipiv, int* d_info, void* stream) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  return getrf_device<float>(m, n, d_a, lda, d_ipiv, d_info, (hipStream_t)stream);
}
int lapack_amd_dgetrs_async(char trans, int n, int nrhs, const double* d_a, int lda, const int* d_ipiv, double* d_b,
                            int ldb, void* stream) {
  int rc = getrs_check<double>(trans, n, nrhs, lda, ldb);
  if (rc) return rc;
  return getrs_device<double>(trans, n, nrhs, d_a, lda, d_ipiv, d_b, ldb, (hipStream_t)stream);
}
int lapack_amd_sgetrs_async(char trans, int n, int nrhs, const float* d_a, int lda, const int* d_ipiv, float* d_b,
                            int ldb, void* stream) {
  int rc = getrs_check<float>(trans, n, nrhs, lda, ldb);
  if (rc) return rc;
  return getrs_device<float>(trans, n, nrhs, d_a, lda, d_ipiv, d_b, ldb, (hipStream_t)stream);
}
int lapack_amd_dlaswp_async(int n, double* d_a, int lda, int k1, int k2, const int* d_ipiv, int incx, void* stream) {
  if (incx != 1 && incx != -1) return LA_ERR_LIMIT;  // strided IPIV: use lapack_amd_dlaswp
  return laswp_device<double>(n, d_a, lda, lda, k1, k2, d_ipiv, incx, (hipStream_t)stream);
}
int lapack_amd_dgemm_async(char ta, char tb, int m, int n, int k, double alpha, const double* d_a, int lda,
                           const double* d_b, int ldb, double beta, double* d_c, int ldc, void* stream) {
  int e = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
  if (e) return -e;
  launch_dgemm(ta, tb, m, n, k, alpha, d_a, lda, d_b, ldb, beta, d_c, ldc, (hipStream_t)stream);
  return hipc(hipGetLastError());
}
int lapack_amd_sgemm_async(char ta, char tb, int m, int n, int k, float alpha, const float* d_a, int lda,
                           const float* d_b, int ldb, float beta, float* d_c, int ldc, void* stream) {
  int e = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
  if (e) return -e;
  launch_sgemm(ta, tb, m, n, k, alpha, d_a, lda, d_b, ldb, beta, d_c, ldc, (hipStream_t)stream);
  return hipc(hipGetLastError());
}

// ---- adjoint of X = A^-1 B -----------------------------------------------------
// Bbar = A^-T Xbar by DGETRS('T') on the saved factors (SRC/dgetrs.f:199-217),
// then Abar = -Bbar X^T (+ Abar if accumulate) as a rank-NRHS product.
int lapack_amd_dgesv_adjoint_async(int n, int nrhs, const double* d_lu, int lda, const int* d_ipiv,
                                   const double* d_x, int ldx, const double* d_xbar, int ldxbar, double* d_bbar,
                                   int ldbbar, double* d_abar, int ldabar, int accumulate, void* stream) {
  if (n < 0) return -1;
  if (nrhs < 0) return -2;
  if (lda < std::max(1, n)) return -4;
  if (ldx < std::max(1, n)) return -7;
  if (ldxbar < std::max(1, n)) return -9;
  if (ldbbar < std::max(1, n)) return -11;
  if (ldabar < std::max(1, n)) return -13;
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (d_bbar != d_xbar) launch_copy_d(n, nrhs, d_xbar, ldxbar, d_bbar, ldbbar, s);
  int rc = getrs_device<double>('T', n, nrhs, d_lu, lda, d_ipiv, d_bbar, ldbbar, s);
  if (rc) return rc;
  const double beta = accumulate ? 1.0 : 0.0;
  if (nrhs <= rankk_max())
    launch_rankk_nt_d(n, n, nrhs, -1.0, d_bbar, ldbbar, d_x, ldx, beta, d_abar, ldabar, s);
  else
    launch_dgemm('N', 'T', n, n, nrhs, -1.0, d_bbar, ldbbar, d_x, ldx, beta, d_abar, ldabar, s);
  return hipc(hipGetLastError());
}

// ---- tuning and tools --------------------------------------------------------
void lapack_amd_set_tuning(int nb, int leaf) { set_tuning(nb, leaf); }
void lapack_amd_get_tuning(int* nb, int* leaf) {
  if (nb) *nb = tune_nb();
  if (leaf) *leaf = tune_leaf();
}
void lapack_amd_phase_enable(int on) { phase_enable(on); }
void lapack_amd_phase_reset(void) { phase_reset(); }
void lapack_amd_phase_read(double* ms, long* calls, double* flops) { phase_read(ms, calls, flops); }
const char* lapack_amd_version(void) { return "lapack_amd 0.2 (gfx950)"; }
int lapack_amd_device_status(int clear) { return device_status(clear); }
int lapack_amd_tune_getrf(int n, const char* path, int* nb, int* leaf, double* ms) {
  return tune_getrf(n, path, nb, leaf, ms);
}
int lapack_amd_load_tuning(const char* path) { return load_tuning(path); }
void lapack_amd_get_tuning_for(int n, int* nb, int* leaf) {
  if (nb) *nb = tune_nb_for(n);
  if (leaf) *leaf = tune_leaf_for(n);
}
void lapack_amd_test_fail_waits(int count) { fail_waits(count); }

}  // extern "C"
