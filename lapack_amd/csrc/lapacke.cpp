// lapacke.cpp -- the LAPACKE C interface (LAPACKE/include/lapacke.h) for the
// routines of the LU solve path, over this library's Fortran drop-ins.
//
// Semantics of the reference's LAPACKE (LAPACKE/src/lapacke_dgetrf.c,
// lapacke_dgetrf_work.c and their siblings): the layout argument is checked
// first (-1); the high-level functions scan their input matrices for NaN
// (LAPACKE_dge_nancheck) unless disabled by LAPACKE_set_nancheck(0) or the
// environment variable LAPACKE_NANCHECK=0, and return minus the argument's
// position; the _work functions call the Fortran routine and shift a
// negative INFO by one (the layout argument comes first); for
// LAPACK_ROW_MAJOR they check the leading dimensions against the row length,
// transpose into column-major temporaries and back.
//
// Arrays may be host or device memory, as for the Fortran drop-ins.  Row-major
// device arrays are transposed on the GPU (a row-major m x n array with
// leading dimension lda is the column-major n x m array A^T); host arrays on
// the host.  NaN scans of device arrays run on the GPU (DLANGE('M') is NaN
// exactly when the matrix holds one).
#include "la_common.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/lapack_amd.h"

namespace la {
struct DevCtx;
DevCtx* ctx_for_current();
hipStream_t ctx_stream(DevCtx* c);
bool is_device_ptr(const void* p);
}  // namespace la

using namespace la;

namespace {

constexpr int ROW = LAPACK_ROW_MAJOR, COL = LAPACK_COL_MAJOR;

int g_nancheck = -1;
std::mutex g_nan_mu;

bool lsame_c(char a, char b) { return (a | 0x20) == (b | 0x20); }

// A row-major array seen as a column-major temporary (or the array itself
// when the layout is column-major).  rows x cols is the logical matrix.
struct Tmp {
  double* user = nullptr;
  int rows = 0, cols = 0, ld = 0;
  bool row = false, dev = false, ok = true;
  std::vector<double> host;
  double* d = nullptr;
  double* p = nullptr;  // what the column-major routine gets
  int ldt = 1;
  hipStream_t s = nullptr;

  Tmp(int layout, double* a, int r, int c, int lda, bool load) : user(a), rows(r), cols(c), ld(lda) {
    row = layout == ROW;
    if (!row || !a) {
      p = a;
      ldt = lda;
      return;
    }
    ldt = std::max(1, r);
    dev = is_device_ptr(a);
    const size_t cnt = (size_t)ldt * std::max(1, c);
    if (dev) {
      DevCtx* cx = ctx_for_current();
      if (!cx || hipMalloc(&d, sizeof(double) * cnt) != hipSuccess) {
        (void)hipGetLastError();
        ok = false;
        return;
      }
      s = ctx_stream(cx);
      p = d;
      // the row-major array is the column-major c x r array A^T
      if (load && r > 0 && c > 0) {
        launch_transpose_d(c, r, a, lda, d, ldt, s);
        ok = hipStreamSynchronize(s) == hipSuccess;
      }
    } else {
      host.assign(cnt, 0.0);
      p = host.data();
      if (load)
        for (int i = 0; i < r; ++i)
          for (int j = 0; j < c; ++j) host[(size_t)i + (size_t)j * ldt] = a[(size_t)i * lda + j];
    }
  }
  void store() {
    if (!row || !user || rows <= 0 || cols <= 0) return;
    if (dev) {
      launch_transpose_d(rows, cols, d, ldt, user, ld, s);
      (void)hipStreamSynchronize(s);
    } else {
      for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) user[(size_t)i * ld + j] = host[(size_t)i + (size_t)j * ldt];
    }
  }
  ~Tmp() {
    if (d) (void)hipFree(d);
  }
};

// LAPACKE_dge_nancheck (LAPACKE/utils/lapacke_dge_nancheck.c): any NaN in the
// logical matrix.
bool ge_has_nan(int layout, int m, int n, const double* a, int lda) {
  if (!a || m <= 0 || n <= 0) return false;
  if (is_device_ptr(a)) {
    // column-major view of the stored rectangle: rows x cols with leading dimension lda
    const int r = layout == COL ? std::min(m, lda) : std::min(n, lda);
    const int c = layout == COL ? n : m;
    return std::isnan(lapack_amd_dlange('M', r, c, a, lda));
  }
  if (layout == COL) {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < std::min(m, lda); ++i)
        if (std::isnan(a[i + (size_t)j * lda])) return true;
  } else {
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < std::min(n, lda); ++j)
        if (std::isnan(a[(size_t)i * lda + j])) return true;
  }
  return false;
}
bool vec_has_nan(int n, const double* x) {
  if (!x) return false;
  std::vector<double> h(std::max(n, 0));
  if (is_device_ptr(x)) {
    if (hipMemcpy(h.data(), x, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) return false;
    x = h.data();
  }
  for (int i = 0; i < n; ++i)
    if (std::isnan(x[i])) return true;
  return false;
}

bool nancheck() { return LAPACKE_get_nancheck() != 0; }
bool bad_layout(int layout, const char* name) {
  if (layout == ROW || layout == COL) return false;
  LAPACKE_xerbla(name, -1);
  return true;
}

}  // namespace

extern "C" {

void LAPACKE_xerbla(const char* name, int info) {
  if (info == LAPACK_WORK_MEMORY_ERROR) std::printf("Not enough memory to allocate work array in %s\n", name);
  else if (info == LAPACK_TRANSPOSE_MEMORY_ERROR) std::printf("Not enough memory to transpose matrix in %s\n", name);
  else if (info < 0) std::printf("Wrong parameter %d in %s\n", -info, name);
}

void LAPACKE_set_nancheck(int flag) {
  std::lock_guard<std::mutex> g(g_nan_mu);
  g_nancheck = flag ? 1 : 0;
}
int LAPACKE_get_nancheck(void) {
  std::lock_guard<std::mutex> g(g_nan_mu);
  if (g_nancheck == -1) {
    const char* e = std::getenv("LAPACKE_NANCHECK");
    g_nancheck = (e == nullptr) ? 1 : (std::atoi(e) ? 1 : 0);
  }
  return g_nancheck;
}

// ---- DGETRF / DGETRF2 --------------------------------------------------------
static int getrf_work(const char* name, int layout, int m, int n, double* a, int lda, int* ipiv, bool rec) {
  int info = 0;
  if (layout == COL) {
    (rec ? dgetrf2_ : dgetrf_)(&m, &n, a, &lda, ipiv, &info);
    return info < 0 ? info - 1 : info;
  }
  if (layout != ROW) { LAPACKE_xerbla(name, -1); return -1; }
  if (lda < n) { LAPACKE_xerbla(name, -5); return -5; }
  Tmp A(layout, a, m, n, lda, true);
  if (!A.ok) { LAPACKE_xerbla(name, LAPACK_TRANSPOSE_MEMORY_ERROR); return LAPACK_TRANSPOSE_MEMORY_ERROR; }
  int ldt = A.ldt;
  (rec ? dgetrf2_ : dgetrf_)(&m, &n, A.p, &ldt, ipiv, &info);
  if (info < 0) info -= 1;
  A.store();
  return info;
}
int LAPACKE_dgetrf_work(int layout, int m, int n, double* a, int lda, int* ipiv) {
  return getrf_work("LAPACKE_dgetrf_work", layout, m, n, a, lda, ipiv, false);
}
int LAPACKE_dgetrf(int layout, int m, int n, double* a, int lda, int* ipiv) {
  if (bad_layout(layout, "LAPACKE_dgetrf")) return -1;
  if (nancheck() && ge_has_nan(layout, m, n, a, lda)) return -4;
  return LAPACKE_dgetrf_work(layout, m, n, a, lda, ipiv);
}
int LAPACKE_dgetrf2_work(int layout, int m, int n, double* a, int lda, int* ipiv) {
  return getrf_work("LAPACKE_dgetrf2_work", layout, m, n, a, lda, ipiv, true);
}
int LAPACKE_dgetrf2(int layout, int m, int n, double* a, int lda, int* ipiv) {
  if (bad_layout(layout, "LAPACKE_dgetrf2")) return -1;
  if (nancheck() && ge_has_nan(layout, m, n, a, lda)) return -4;
  return LAPACKE_dgetrf2_work(layout, m, n, a, lda, ipiv);
}

// ---- DGETRS ------------------------------------------------------------------
int LAPACKE_dgetrs_work(int layout, char trans, int n, int nrhs, const double* a, int lda, const int* ipiv,
                        double* b, int ldb) {
  int info = 0;
  if (layout == COL) {
    dgetrs_(&trans, &n, &nrhs, a, &lda, ipiv, b, &ldb, &info, 1);
    return info < 0 ? info - 1 : info;
  }
  if (layout != ROW) { LAPACKE_xerbla("LAPACKE_dgetrs_work", -1); return -1; }
  if (lda < n) { LAPACKE_xerbla("LAPACKE_dgetrs_work", -6); return -6; }
  if (ldb < nrhs) { LAPACKE_xerbla("LAPACKE_dgetrs_work", -9); return -9; }
  Tmp A(layout, const_cast<double*>(a), n, n, lda, true);
  Tmp B(layout, b, n, nrhs, ldb, true);
  if (!A.ok || !B.ok) {
    LAPACKE_xerbla("LAPACKE_dgetrs_work", LAPACK_TRANSPOSE_MEMORY_ERROR);
    return LAPACK_TRANSPOSE_MEMORY_ERROR;
  }
  int lda_t = A.ldt, ldb_t = B.ldt;
  dgetrs_(&trans, &n, &nrhs, A.p, &lda_t, ipiv, B.p, &ldb_t, &info, 1);
  if (info < 0) info -= 1;
  B.store();
  return info;
}
int LAPACKE_dgetrs(int layout, char trans, int n, int nrhs, const double* a, int lda, const int* ipiv, double* b,
                   int ldb) {
  if (bad_layout(layout, "LAPACKE_dgetrs")) return -1;
  if (nancheck()) {
    if (ge_has_nan(layout, n, n, a, lda)) return -5;
    if (ge_has_nan(layout, n, nrhs, b, ldb)) return -8;
  }
  return LAPACKE_dgetrs_work(layout, trans, n, nrhs, a, lda, ipiv, b, ldb);
}

// ---- DGESV -------------------------------------------------------------------
int LAPACKE_dgesv_work(int layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb) {
  int info = 0;
  if (layout == COL) {
    dgesv_(&n, &nrhs, a, &lda, ipiv, b, &ldb, &info);
    return info < 0 ? info - 1 : info;
  }
  if (layout != ROW) { LAPACKE_xerbla("LAPACKE_dgesv_work", -1); return -1; }
  if (lda < n) { LAPACKE_xerbla("LAPACKE_dgesv_work", -5); return -5; }
  if (ldb < nrhs) { LAPACKE_xerbla("LAPACKE_dgesv_work", -8); return -8; }
  Tmp A(layout, a, n, n, lda, true);
  Tmp B(layout, b, n, nrhs, ldb, true);
  if (!A.ok || !B.ok) {
    LAPACKE_xerbla("LAPACKE_dgesv_work", LAPACK_TRANSPOSE_MEMORY_ERROR);
    return LAPACK_TRANSPOSE_MEMORY_ERROR;
  }
  int lda_t = A.ldt, ldb_t = B.ldt;
  dgesv_(&n, &nrhs, A.p, &lda_t, ipiv, B.p, &ldb_t, &info);
  if (info < 0) info -= 1;
  A.store();
  B.store();
  return info;
}
int LAPACKE_dgesv(int layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb) {
  if (bad_layout(layout, "LAPACKE_dgesv")) return -1;
  if (nancheck()) {
    if (ge_has_nan(layout, n, n, a, lda)) return -4;
    if (ge_has_nan(layout, n, nrhs, b, ldb)) return -7;
  }
  return LAPACKE_dgesv_work(layout, n, nrhs, a, lda, ipiv, b, ldb);
}

// ---- DGETRI ------------------------------------------------------------------
int LAPACKE_dgetri_work(int layout, int n, double* a, int lda, const int* ipiv, double* work, int lwork) {
  int info = 0;
  if (layout == COL) {
    dgetri_(&n, a, &lda, ipiv, work, &lwork, &info);
    return info < 0 ? info - 1 : info;
  }
  if (layout != ROW) { LAPACKE_xerbla("LAPACKE_dgetri_work", -1); return -1; }
  if (lda < n) { LAPACKE_xerbla("LAPACKE_dgetri_work", -4); return -4; }
  int lda_t = std::max(1, n);
  if (lwork == -1) {  // workspace query
    dgetri_(&n, a, &lda_t, ipiv, work, &lwork, &info);
    return info < 0 ? info - 1 : info;
  }
  Tmp A(layout, a, n, n, lda, true);
  if (!A.ok) {
    LAPACKE_xerbla("LAPACKE_dgetri_work", LAPACK_TRANSPOSE_MEMORY_ERROR);
    return LAPACK_TRANSPOSE_MEMORY_ERROR;
  }
  dgetri_(&n, A.p, &lda_t, ipiv, work, &lwork, &info);
  if (info < 0) info -= 1;
  A.store();
  return info;
}
int LAPACKE_dgetri(int layout, int n, double* a, int lda, const int* ipiv) {
  if (bad_layout(layout, "LAPACKE_dgetri")) return -1;
  if (nancheck() && ge_has_nan(layout, n, n, a, lda)) return -3;
  double q = 0.0;
  int info = LAPACKE_dgetri_work(layout, n, a, lda, ipiv, &q, -1);
  if (info != 0) return info;
  std::vector<double> work(std::max(1, (int)q));
  return LAPACKE_dgetri_work(layout, n, a, lda, ipiv, work.data(), (int)work.size());
}

// ---- DSGESV ------------------------------------------------------------------
int LAPACKE_dsgesv_work(int layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb, double* x,
                        int ldx, double* work, float* swork, int* iter) {
  int info = 0;
  if (layout == COL) {
    dsgesv_(&n, &nrhs, a, &lda, ipiv, b, &ldb, x, &ldx, work, swork, iter, &info);
    return info < 0 ? info - 1 : info;
  }
  if (layout != ROW) { LAPACKE_xerbla("LAPACKE_dsgesv_work", -1); return -1; }
  if (lda < n) { LAPACKE_xerbla("LAPACKE_dsgesv_work", -5); return -5; }
  if (ldb < nrhs) { LAPACKE_xerbla("LAPACKE_dsgesv_work", -8); return -8; }
  if (ldx < nrhs) { LAPACKE_xerbla("LAPACKE_dsgesv_work", -10); return -10; }
  Tmp A(layout, a, n, n, lda, true);
  Tmp B(layout, b, n, nrhs, ldb, true);
  Tmp X(layout, x, n, nrhs, ldx, false);
  if (!A.ok || !B.ok || !X.ok) {
    LAPACKE_xerbla("LAPACKE_dsgesv_work", LAPACK_TRANSPOSE_MEMORY_ERROR);
    return LAPACK_TRANSPOSE_MEMORY_ERROR;
  }
  int lda_t = A.ldt, ldb_t = B.ldt, ldx_t = X.ldt;
  dsgesv_(&n, &nrhs, A.p, &lda_t, ipiv, B.p, &ldb_t, X.p, &ldx_t, work, swork, iter, &info);
  if (info < 0) info -= 1;
  A.store();
  X.store();
  return info;
}
int LAPACKE_dsgesv(int layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb, double* x, int ldx,
                   int* iter) {
  if (bad_layout(layout, "LAPACKE_dsgesv")) return -1;
  if (nancheck()) {
    if (ge_has_nan(layout, n, n, a, lda)) return -4;
    if (ge_has_nan(layout, n, nrhs, b, ldb)) return -7;
  }
  // WORK and SWORK are not used by the GPU solver (it keeps its own device
  // workspace); the reference allocates them here, so a failure to allocate
  // is not a case that arises
  return LAPACKE_dsgesv_work(layout, n, nrhs, a, lda, ipiv, b, ldb, x, ldx, nullptr, nullptr, iter);
}

// ---- DGECON ------------------------------------------------------------------
int LAPACKE_dgecon_work(int layout, char norm, int n, const double* a, int lda, double anorm, double* rcond,
                        double* work, int* iwork) {
  int info = 0;
  if (layout == COL) {
    dgecon_(&norm, &n, a, &lda, &anorm, rcond, work, iwork, &info, 1);
    return info < 0 ? info - 1 : info;
  }
  if (layout != ROW) { LAPACKE_xerbla("LAPACKE_dgecon_work", -1); return -1; }
  if (lda < n) { LAPACKE_xerbla("LAPACKE_dgecon_work", -5); return -5; }
  Tmp A(layout, const_cast<double*>(a), n, n, lda, true);
  if (!A.ok) {
    LAPACKE_xerbla("LAPACKE_dgecon_work", LAPACK_TRANSPOSE_MEMORY_ERROR);
    return LAPACK_TRANSPOSE_MEMORY_ERROR;
  }
  int lda_t = A.ldt;
  dgecon_(&norm, &n, A.p, &lda_t, &anorm, rcond, work, iwork, &info, 1);
  return info < 0 ? info - 1 : info;
}
int LAPACKE_dgecon(int layout, char norm, int n, const double* a, int lda, double anorm, double* rcond) {
  if (bad_layout(layout, "LAPACKE_dgecon")) return -1;
  if (nancheck()) {
    if (ge_has_nan(layout, n, n, a, lda)) return -4;
    if (std::isnan(anorm)) return -6;
  }
  return LAPACKE_dgecon_work(layout, norm, n, a, lda, anorm, rcond, nullptr, nullptr);
}

// ---- DLANGE ------------------------------------------------------------------
double LAPACKE_dlange_work(int layout, char norm, int m, int n, const double* a, int lda, double* work) {
  if (layout == COL) return dlange_(&norm, &m, &n, a, &lda, work, 1);
  if (layout != ROW) { LAPACKE_xerbla("LAPACKE_dlange_work", -1); return -1; }
  if (lda < n) { LAPACKE_xerbla("LAPACKE_dlange_work", -6); return -6; }
  // the row-major array is the column-major n x m transpose: swap the one
  // and infinity norms instead of transposing
  char nl = norm;
  if (lsame_c(norm, '1') || lsame_c(norm, 'o')) nl = 'i';
  else if (lsame_c(norm, 'i')) nl = '1';
  return dlange_(&nl, &n, &m, a, &lda, work, 1);
}
double LAPACKE_dlange(int layout, char norm, int m, int n, const double* a, int lda) {
  if (bad_layout(layout, "LAPACKE_dlange")) return -1;
  if (nancheck() && ge_has_nan(layout, m, n, a, lda)) return -5;
  return LAPACKE_dlange_work(layout, norm, m, n, a, lda, nullptr);
}

// ---- DLASWP ------------------------------------------------------------------
int LAPACKE_dlaswp_work(int layout, int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  if (layout == COL) {
    dlaswp_(&n, a, &lda, &k1, &k2, ipiv, &incx);
    return 0;
  }
  if (layout != ROW) { LAPACKE_xerbla("LAPACKE_dlaswp_work", -1); return -1; }
  if (lda < n) { LAPACKE_xerbla("LAPACKE_dlaswp_work", -4); return -4; }
  // rows reached by the interchanges: k2 and every target (IPIV is read on the host)
  std::vector<int> hp(std::max(0, k1 + (k2 - k1) * std::abs(incx)));
  const int* ip = ipiv;
  if (is_device_ptr(ipiv) && !hp.empty()) {
    if (hipMemcpy(hp.data(), ipiv, sizeof(int) * hp.size(), hipMemcpyDeviceToHost) != hipSuccess) return -7;
    ip = hp.data();
  }
  int rows = std::max(1, k2);
  for (int i = k1; i <= k2; ++i) rows = std::max(rows, ip[k1 + (i - k1) * std::abs(incx) - 1]);
  Tmp A(layout, a, rows, n, lda, true);
  if (!A.ok) {
    LAPACKE_xerbla("LAPACKE_dlaswp_work", LAPACK_TRANSPOSE_MEMORY_ERROR);
    return LAPACK_TRANSPOSE_MEMORY_ERROR;
  }
  int lda_t = A.ldt;
  dlaswp_(&n, A.p, &lda_t, &k1, &k2, ipiv, &incx);
  A.store();
  return 0;
}
int LAPACKE_dlaswp(int layout, int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  if (bad_layout(layout, "LAPACKE_dlaswp")) return -1;
  // the reference disables this routine's NaN check (LAPACKE/src/lapacke_dlaswp.c)
  return LAPACKE_dlaswp_work(layout, n, a, lda, k1, k2, ipiv, incx);
}

// ---- DGESVX ------------------------------------------------------------------
int LAPACKE_dgesvx_work(int layout, char fact, char trans, int n, int nrhs, double* a, int lda, double* af, int ldaf,
                        int* ipiv, char* equed, double* r, double* c, double* b, int ldb, double* x, int ldx,
                        double* rcond, double* ferr, double* berr, double* work, int* iwork) {
  int info = 0;
  if (layout == COL) {
    dgesvx_(&fact, &trans, &n, &nrhs, a, &lda, af, &ldaf, ipiv, equed, r, c, b, &ldb, x, &ldx, rcond, ferr, berr,
            work, iwork, &info, 1, 1, 1);
    return info < 0 ? info - 1 : info;
  }
  if (layout != ROW) { LAPACKE_xerbla("LAPACKE_dgesvx_work", -1); return -1; }
  if (lda < n) { LAPACKE_xerbla("LAPACKE_dgesvx_work", -7); return -7; }
  if (ldaf < n) { LAPACKE_xerbla("LAPACKE_dgesvx_work", -9); return -9; }
  if (ldb < nrhs) { LAPACKE_xerbla("LAPACKE_dgesvx_work", -15); return -15; }
  if (ldx < nrhs) { LAPACKE_xerbla("LAPACKE_dgesvx_work", -17); return -17; }
  Tmp A(layout, a, n, n, lda, true);
  Tmp F(layout, af, n, n, ldaf, lsame_c(fact, 'f'));
  Tmp B(layout, b, n, nrhs, ldb, true);
  Tmp X(layout, x, n, nrhs, ldx, false);
  if (!A.ok || !F.ok || !B.ok || !X.ok) {
    LAPACKE_xerbla("LAPACKE_dgesvx_work", LAPACK_TRANSPOSE_MEMORY_ERROR);
    return LAPACK_TRANSPOSE_MEMORY_ERROR;
  }
  int lda_t = A.ldt, ldaf_t = F.ldt, ldb_t = B.ldt, ldx_t = X.ldt;
  dgesvx_(&fact, &trans, &n, &nrhs, A.p, &lda_t, F.p, &ldaf_t, ipiv, equed, r, c, B.p, &ldb_t, X.p, &ldx_t, rcond,
          ferr, berr, work, iwork, &info, 1, 1, 1);
  if (info < 0) info -= 1;
  if (lsame_c(fact, 'e') && (*equed == 'B' || *equed == 'C' || *equed == 'R')) A.store();
  if (lsame_c(fact, 'e') || lsame_c(fact, 'n')) F.store();
  B.store();
  X.store();
  return info;
}
int LAPACKE_dgesvx(int layout, char fact, char trans, int n, int nrhs, double* a, int lda, double* af, int ldaf,
                   int* ipiv, char* equed, double* r, double* c, double* b, int ldb, double* x, int ldx, double* rcond,
                   double* ferr, double* berr, double* rpivot) {
  if (bad_layout(layout, "LAPACKE_dgesvx")) return -1;
  if (nancheck()) {
    if (ge_has_nan(layout, n, n, a, lda)) return -6;
    if (lsame_c(fact, 'f') && ge_has_nan(layout, n, n, af, ldaf)) return -8;
    if (ge_has_nan(layout, n, nrhs, b, ldb)) return -14;
    if (lsame_c(fact, 'f') && (lsame_c(*equed, 'b') || lsame_c(*equed, 'c')) && vec_has_nan(n, c)) return -13;
    if (lsame_c(fact, 'f') && (lsame_c(*equed, 'b') || lsame_c(*equed, 'r')) && vec_has_nan(n, r)) return -12;
  }
  std::vector<double> work(std::max(1, 4 * n));
  std::vector<int> iwork(std::max(1, n));
  const int info = LAPACKE_dgesvx_work(layout, fact, trans, n, nrhs, a, lda, af, ldaf, ipiv, equed, r, c, b, ldb, x,
                                       ldx, rcond, ferr, berr, work.data(), iwork.data());
  if (rpivot) *rpivot = work[0];
  return info;
}

}  // extern "C"
