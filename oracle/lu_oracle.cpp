// lu_oracle.cpp -- CPU restatement of Reference-LAPACK's dense LU solve path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file; it is the checker that the GPU path is compared against
// (tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py).
//
// Every routine follows the loop order of the reference routine it names so
// that, compiled with `-O2 -ffp-contract=off` (no FMA contraction, no
// -march=native), it is bit-identical to the Fortran reference built from
// its own sources (oracle/Makefile -> oracle/_ref).  Parity of this file with
// that build is asserted by tests/test_oracle.py on a size sweep.
//
// Index conventions: arrays are column-major with leading dimension ld; the
// routines use Fortran's 1-based row/column numbers internally (AT(i,j)) and
// keep IPIV 1-based exactly as the reference does.
//
// Pinning (how each function is checked, see tests/test_oracle.py):
//   iamax, scal, laswp, trsm_left, gemm, getrf2, getrf, getrs  -- bitwise
//     equality with the flang-built reference (oracle/_ref) on seeded inputs;
//     LU also by the DGET01 ratio, x by DGET02, and brute force on tiny cases.
//   lange, get01/get02/get08                                   -- closed forms
//     on hand-built matrices (tests/test_oracle.py) and agreement with _ref.
//   lag2s / lag2d / dsgesv                                     -- bitwise
//     equality with _ref's DSGESV (ITER, IPIV, X).
//   geequ / geequb / laqge                                     -- DCHKEQ
//     closed form (TESTING/LIN/dchkeq.f:107-173) and equality with _ref.
//   gemv, asum, lacn2, gerfs                                   -- bitwise
//     equality with _ref's DGERFS (X, FERR, BERR) and DGET07's bounds.
//   get04 / get07                                              -- equality
//     with the reference's TESTING/LIN/dget04.f and dget07.f built into _ref,
//     and closed forms (known errors, FERR too small, BERR at its bound).
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <cfloat>
#include <vector>
#include <algorithm>

namespace {

template <class T> struct Mach;
// DLAMCH('E') = EPSILON(0d0)*0.5 (INSTALL/dlamch.f:101-107); 'S' = TINY since
// 1/HUGE < TINY (INSTALL/dlamch.f:111-123); 'O' = HUGE.
template <> struct Mach<double> {
  static constexpr double eps = 0x1p-53;
  static constexpr double sfmin = 0x1p-1022;
  static constexpr double ovfl = DBL_MAX;
};
// SLAMCH, same definitions in single precision (INSTALL/slamch.f).
template <> struct Mach<float> {
  static constexpr float eps = 0x1p-24f;
  static constexpr float sfmin = 0x1p-126f;
  static constexpr float ovfl = FLT_MAX;
};

template <class T> inline T tabs(T x) { return x < T(0) ? -x : (x == T(0) ? T(0) : x); }

// Column-major element (i,j), 1-based, leading dimension ld.
#define AT(a, i, j, ld) (a)[(size_t)((i)-1) + (size_t)((j)-1) * (size_t)(ld)]

// IDAMAX: first index of max |x|; strict '>' so ties keep the lowest index
// and a NaN is only ever returned at position 1 (BLAS/SRC/idamax.f:93-122).
template <class T> int iamax(int n, const T* x, int incx) {
  if (n < 1 || incx <= 0) return 0;
  if (n == 1) return 1;
  int imax = 1;
  T dmax = std::fabs(x[0]);
  size_t ix = (size_t)incx;
  for (int i = 2; i <= n; ++i, ix += (size_t)incx) {
    T v = std::fabs(x[ix]);
    if (v > dmax) { imax = i; dmax = v; }
  }
  return imax;
}

// DSCAL x <- da*x; quick return when da == 1 (BLAS/SRC/dscal.f:104-135).
template <class T> void scal(int n, T da, T* x, int incx) {
  if (n <= 0 || incx <= 0 || da == T(1)) return;
  for (int i = 0; i < n; ++i) x[(size_t)i * incx] = da * x[(size_t)i * incx];
}

// DLASWP: sequential row interchanges row I <-> row IPIV(IX) for I = K1..K2
// (INCX > 0) or K2..K1 (INCX < 0) (SRC/dlaswp.f:138-183).  The reference
// sweeps 32-column blocks; the swaps of different columns are independent,
// so applying the whole sequence column block by column block or row by row
// moves the same values.
template <class T> void laswp(int n, T* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  int ix0, i1, i2, inc;
  if (incx > 0) { ix0 = k1; i1 = k1; i2 = k2; inc = 1; }
  else if (incx < 0) { ix0 = k1 + (k1 - k2) * incx; i1 = k2; i2 = k1; inc = -1; }
  else return;
  int ix = ix0;
  for (int i = i1; inc > 0 ? i <= i2 : i >= i2; i += inc) {
    int ip = ipiv[ix - 1];
    if (ip != i) {
      for (int k = 1; k <= n; ++k) std::swap(AT(a, i, k, lda), AT(a, ip, k, lda));
    }
    ix += incx;
  }
}

// DTRSM with SIDE='L' (BLAS/SRC/dtrsm.f:261-329): B <- alpha*op(A)^-1*B.
template <class T>
void trsm_left(char uplo, char trans, char diag, int m, int n, T alpha, const T* a, int lda,
               T* b, int ldb) {
  if (m == 0 || n == 0) return;
  const bool upper = (uplo == 'U' || uplo == 'u');
  const bool notran = (trans == 'N' || trans == 'n');
  const bool nounit = (diag == 'N' || diag == 'n');
  if (alpha == T(0)) {
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m; ++i) AT(b, i, j, ldb) = T(0);
    return;
  }
  if (notran) {
    if (upper) {  // dtrsm.f:277-288
      for (int j = 1; j <= n; ++j) {
        for (int i = 1; i <= m; ++i) AT(b, i, j, ldb) = alpha * AT(b, i, j, ldb);
        for (int k = m; k >= 1; --k) {
          if (nounit) AT(b, k, j, ldb) = AT(b, k, j, ldb) / AT(a, k, k, lda);
          const T bkj = AT(b, k, j, ldb);
          for (int i = 1; i <= k - 1; ++i) AT(b, i, j, ldb) = AT(b, i, j, ldb) - bkj * AT(a, i, k, lda);
        }
      }
    } else {  // dtrsm.f:289-300
      for (int j = 1; j <= n; ++j) {
        for (int i = 1; i <= m; ++i) AT(b, i, j, ldb) = alpha * AT(b, i, j, ldb);
        for (int k = 1; k <= m; ++k) {
          if (nounit) AT(b, k, j, ldb) = AT(b, k, j, ldb) / AT(a, k, k, lda);
          const T bkj = AT(b, k, j, ldb);
          for (int i = k + 1; i <= m; ++i) AT(b, i, j, ldb) = AT(b, i, j, ldb) - bkj * AT(a, i, k, lda);
        }
      }
    }
  } else {
    if (upper) {  // dtrsm.f:306-316
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= m; ++i) {
          T temp = alpha * AT(b, i, j, ldb);
          for (int k = 1; k <= i - 1; ++k) temp = temp - AT(a, k, i, lda) * AT(b, k, j, ldb);
          if (nounit) temp = temp / AT(a, i, i, lda);
          AT(b, i, j, ldb) = temp;
        }
    } else {  // dtrsm.f:317-327
      for (int j = 1; j <= n; ++j)
        for (int i = m; i >= 1; --i) {
          T temp = alpha * AT(b, i, j, ldb);
          for (int k = i + 1; k <= m; ++k) temp = temp - AT(a, k, i, lda) * AT(b, k, j, ldb);
          if (nounit) temp = temp / AT(a, i, i, lda);
          AT(b, i, j, ldb) = temp;
        }
    }
  }
}

// DGEMM C <- alpha*op(A)*op(B) + beta*C (BLAS/SRC/dgemm.f:298-402).
template <class T>
void gemm(char transa, char transb, int m, int n, int k, T alpha, const T* a, int lda, const T* b,
          int ldb, T beta, T* c, int ldc) {
  const bool nota = (transa == 'N' || transa == 'n');
  const bool notb = (transb == 'N' || transb == 'n');
  if (m == 0 || n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return;
  if (alpha == T(0)) {
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m; ++i)
        AT(c, i, j, ldc) = (beta == T(0)) ? T(0) : beta * AT(c, i, j, ldc);
    return;
  }
  if (notb) {
    if (nota) {  // dgemm.f:327-343
      for (int j = 1; j <= n; ++j) {
        if (beta == T(0)) { for (int i = 1; i <= m; ++i) AT(c, i, j, ldc) = T(0); }
        else if (beta != T(1)) { for (int i = 1; i <= m; ++i) AT(c, i, j, ldc) = beta * AT(c, i, j, ldc); }
        for (int l = 1; l <= k; ++l) {
          const T temp = alpha * AT(b, l, j, ldb);
          T* cj = &AT(c, 1, j, ldc);
          const T* al = &AT(a, 1, l, lda);
          for (int i = 0; i < m; ++i) cj[i] = cj[i] + temp * al[i];
        }
      }
    } else {  // dgemm.f:348-360
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= m; ++i) {
          T temp = T(0);
          for (int l = 1; l <= k; ++l) temp = temp + AT(a, l, i, lda) * AT(b, l, j, ldb);
          AT(c, i, j, ldc) = (beta == T(0)) ? alpha * temp : alpha * temp + beta * AT(c, i, j, ldc);
        }
    }
  } else {
    if (nota) {  // dgemm.f:367-383
      for (int j = 1; j <= n; ++j) {
        if (beta == T(0)) { for (int i = 1; i <= m; ++i) AT(c, i, j, ldc) = T(0); }
        else if (beta != T(1)) { for (int i = 1; i <= m; ++i) AT(c, i, j, ldc) = beta * AT(c, i, j, ldc); }
        for (int l = 1; l <= k; ++l) {
          const T temp = alpha * AT(b, j, l, ldb);
          T* cj = &AT(c, 1, j, ldc);
          const T* al = &AT(a, 1, l, lda);
          for (int i = 0; i < m; ++i) cj[i] = cj[i] + temp * al[i];
        }
      }
    } else {  // dgemm.f:388-400
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= m; ++i) {
          T temp = T(0);
          for (int l = 1; l <= k; ++l) temp = temp + AT(a, l, i, lda) * AT(b, j, l, ldb);
          AT(c, i, j, ldc) = (beta == T(0)) ? alpha * temp : alpha * temp + beta * AT(c, i, j, ldc);
        }
    }
  }
}

// DGETRF2: recursive panel LU (SRC/dgetrf2.f:152-265).
template <class T> int getrf2(int m, int n, T* a, int lda, int* ipiv) {
  int info = 0;
  if (m == 0 || n == 0) return 0;
  if (m == 1) {  // dgetrf2.f:170-177
    ipiv[0] = 1;
    if (AT(a, 1, 1, lda) == T(0)) info = 1;
  } else if (n == 1) {  // dgetrf2.f:179-214
    const T sfmin = Mach<T>::sfmin;
    const int i = iamax(m, a, 1);
    ipiv[0] = i;
    if (AT(a, i, 1, lda) != T(0)) {
      if (i != 1) std::swap(AT(a, 1, 1, lda), AT(a, i, 1, lda));
      if (std::fabs(AT(a, 1, 1, lda)) >= sfmin) {
        scal(m - 1, T(1) / AT(a, 1, 1, lda), &AT(a, 2, 1, lda), 1);
      } else {
        for (int r = 1; r <= m - 1; ++r) AT(a, 1 + r, 1, lda) = AT(a, 1 + r, 1, lda) / AT(a, 1, 1, lda);
      }
    } else {
      info = 1;
    }
  } else {  // dgetrf2.f:220-263
    const int n1 = std::min(m, n) / 2;
    const int n2 = n - n1;
    int iinfo = getrf2(m, n1, a, lda, ipiv);
    if (info == 0 && iinfo > 0) info = iinfo;
    laswp(n2, &AT(a, 1, n1 + 1, lda), lda, 1, n1, ipiv, 1);
    trsm_left('L', 'N', 'U', n1, n2, T(1), a, lda, &AT(a, 1, n1 + 1, lda), lda);
    gemm('N', 'N', m - n1, n2, n1, T(-1), &AT(a, n1 + 1, 1, lda), lda, &AT(a, 1, n1 + 1, lda), lda,
         T(1), &AT(a, n1 + 1, n1 + 1, lda), lda);
    iinfo = getrf2(m - n1, n2, &AT(a, n1 + 1, n1 + 1, lda), lda, ipiv + n1);
    if (info == 0 && iinfo > 0) info = iinfo + n1;
    for (int i = n1 + 1; i <= std::min(m, n); ++i) ipiv[i - 1] += n1;
    laswp(n1, a, lda, n1 + 1, std::min(m, n), ipiv, 1);
  }
  return info;
}

// DGETRF: blocked right-looking LU (SRC/dgetrf.f:144-220); nb plays the role
// of ILAENV(1,'DGETRF',...) which is 64 in the reference (SRC/ilaenv.f:289-295).
template <class T> int getrf(int m, int n, T* a, int lda, int* ipiv, int nb) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  if (m == 0 || n == 0) return 0;
  int info = 0;
  if (nb <= 1 || nb >= std::min(m, n)) return getrf2(m, n, a, lda, ipiv);
  for (int j = 1; j <= std::min(m, n); j += nb) {
    const int jb = std::min(std::min(m, n) - j + 1, nb);
    const int iinfo = getrf2(m - j + 1, jb, &AT(a, j, j, lda), lda, ipiv + (j - 1));
    if (info == 0 && iinfo > 0) info = iinfo + j - 1;
    for (int i = j; i <= std::min(m, j + jb - 1); ++i) ipiv[i - 1] = j - 1 + ipiv[i - 1];
    laswp(j - 1, a, lda, j, j + jb - 1, ipiv, 1);
    if (j + jb <= n) {
      laswp(n - j - jb + 1, &AT(a, 1, j + jb, lda), lda, j, j + jb - 1, ipiv, 1);
      trsm_left('L', 'N', 'U', jb, n - j - jb + 1, T(1), &AT(a, j, j, lda), lda, &AT(a, j, j + jb, lda), lda);
      if (j + jb <= m)
        gemm('N', 'N', m - j - jb + 1, n - j - jb + 1, jb, T(-1), &AT(a, j + jb, j, lda), lda,
             &AT(a, j, j + jb, lda), lda, T(1), &AT(a, j + jb, j + jb, lda), lda);
    }
  }
  return info;
}

// DGETRS (SRC/dgetrs.f:157-218).
template <class T>
int getrs(char trans, int n, int nrhs, const T* a, int lda, const int* ipiv, T* b, int ldb) {
  const bool notran = (trans == 'N' || trans == 'n');
  if (!notran && !(trans == 'T' || trans == 't' || trans == 'C' || trans == 'c')) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max(1, n)) return -5;
  if (ldb < std::max(1, n)) return -8;
  if (n == 0 || nrhs == 0) return 0;
  if (notran) {
    laswp(nrhs, b, ldb, 1, n, ipiv, 1);
    trsm_left('L', 'N', 'U', n, nrhs, T(1), a, lda, b, ldb);
    trsm_left('U', 'N', 'N', n, nrhs, T(1), a, lda, b, ldb);
  } else {
    trsm_left('U', 'T', 'N', n, nrhs, T(1), a, lda, b, ldb);
    trsm_left('L', 'T', 'U', n, nrhs, T(1), a, lda, b, ldb);
    laswp(nrhs, b, ldb, 1, n, ipiv, -1);
  }
  return 0;
}

// DLANGE for NORM in {'M','1'/'O','I'} (SRC/dlange.f:115-185).
// DLASSQ (SRC/dlassq.f90): scale and sumsq with scale^2*sumsq updated by
// sum(x_i^2), Blue's three accumulators; constants of SRC/la_constants.f90
// for FP64 (radix 2, minexponent -1021, maxexponent 1024, digits 53):
// tsml = 2^-511, tbig = 2^486, ssml = 2^537, sbig = 2^-538.
inline void lassq(int n, const double* x, double& scale, double& sumsq) {
  const double tsml = 0x1p-511, tbig = 0x1p486, ssml = 0x1p537, sbig = 0x1p-538;
  if (std::isnan(scale) || std::isnan(sumsq)) return;
  if (sumsq == 0.0) scale = 1.0;
  if (scale == 0.0) {
    scale = 1.0;
    sumsq = 0.0;
  }
  if (n <= 0) return;
  bool notbig = true;
  double asml = 0.0, amed = 0.0, abig = 0.0;
  for (int i = 0; i < n; ++i) {
    const double ax = std::fabs(x[i]);
    if (ax > tbig) {
      abig = abig + (ax * sbig) * (ax * sbig);
      notbig = false;
    } else if (ax < tsml) {
      if (notbig) asml = asml + (ax * ssml) * (ax * ssml);
    } else {
      amed = amed + ax * ax;
    }
  }
  if (sumsq > 0.0) {
    const double ax = scale * std::sqrt(sumsq);
    if (ax > tbig) {
      if (scale > 1.0) {
        scale = scale * sbig;
        abig = abig + scale * (scale * sumsq);
      } else {
        abig = abig + scale * (scale * (sbig * (sbig * sumsq)));
      }
    } else if (ax < tsml) {
      if (notbig) {
        if (scale < 1.0) {
          scale = scale * ssml;
          asml = asml + scale * (scale * sumsq);
        } else {
          asml = asml + scale * (scale * (ssml * (ssml * sumsq)));
        }
      }
    } else {
      amed = amed + scale * (scale * sumsq);
    }
  }
  if (abig > 0.0) {
    if (amed > 0.0 || std::isnan(amed)) abig = abig + (amed * sbig) * sbig;
    scale = 1.0 / sbig;
    sumsq = abig;
  } else if (asml > 0.0) {
    if (amed > 0.0 || std::isnan(amed)) {
      amed = std::sqrt(amed);
      asml = std::sqrt(asml) / ssml;
      const double ymin = asml > amed ? amed : asml;
      const double ymax = asml > amed ? asml : amed;
      scale = 1.0;
      sumsq = ymax * ymax * (1.0 + (ymin / ymax) * (ymin / ymax));
    } else {
      scale = 1.0 / ssml;
      sumsq = asml;
    }
  } else {
    scale = 1.0;
    sumsq = amed;
  }
}

template <class T> T lange(char norm, int m, int n, const T* a, int lda) {
  if (std::min(m, n) == 0) return T(0);
  T value = T(0);
  if (norm == 'M' || norm == 'm') {
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m; ++i) {
        T temp = std::fabs(AT(a, i, j, lda));
        if (value < temp || std::isnan(temp)) value = temp;
      }
  } else if (norm == 'O' || norm == 'o' || norm == '1') {
    for (int j = 1; j <= n; ++j) {
      T sum = T(0);
      for (int i = 1; i <= m; ++i) sum = sum + std::fabs(AT(a, i, j, lda));
      if (value < sum || std::isnan(sum)) value = sum;
    }
  } else if (norm == 'I' || norm == 'i') {
    std::vector<T> work((size_t)m, T(0));
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m; ++i) work[i - 1] = work[i - 1] + std::fabs(AT(a, i, j, lda));
    for (int i = 0; i < m; ++i) {
      T temp = work[i];
      if (value < temp || std::isnan(temp)) value = temp;
    }
  } else if (norm == 'F' || norm == 'f' || norm == 'E' || norm == 'e') {
    // dlange.f:190-200: DLASSQ over the columns
    double scale = 0.0, sum = 1.0;
    std::vector<double> col((size_t)m);
    for (int j = 1; j <= n; ++j) {
      for (int i = 1; i <= m; ++i) col[i - 1] = (double)AT(a, i, j, lda);
      lassq(m, col.data(), scale, sum);
    }
    value = (T)(scale * std::sqrt(sum));
  }
  return value;
}

// DLAG2S: FP64 -> FP32 by round to nearest; INFO=1 on |a| > SLAMCH('O')
// (SRC/dlag2s.f:135-146).
int lag2s(int m, int n, const double* a, int lda, float* sa, int ldsa) {
  const double rmax = (double)FLT_MAX;
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= m; ++i) {
      const double v = AT(a, i, j, lda);
      if (v < -rmax || v > rmax) return 1;
      AT(sa, i, j, ldsa) = (float)v;
    }
  return 0;
}

// SLAG2D: FP32 -> FP64, exact (SRC/slag2d.f:122-128).
void lag2d(int m, int n, const float* sa, int ldsa, double* a, int lda) {
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= m; ++i) AT(a, i, j, lda) = (double)AT(sa, i, j, ldsa);
}

// DGEMV y <- alpha*op(A)*x + beta*y, incx = incy = 1 (BLAS/SRC/dgemv.f:254-330):
// 'N' as column axpys, 'T' as column dot products.
void gemv(char trans, int m, int n, double alpha, const double* a, int lda, const double* x, double beta,
          double* y) {
  const bool nt = (trans == 'N' || trans == 'n');
  const int leny = nt ? m : n, lenx = nt ? n : m;
  if (m == 0 || n == 0 || (alpha == 0.0 && beta == 1.0)) return;
  if (beta != 1.0) {
    for (int i = 0; i < leny; ++i) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
  }
  if (alpha == 0.0) return;
  (void)lenx;
  if (nt) {
    for (int j = 1; j <= n; ++j) {
      const double temp = alpha * x[j - 1];
      for (int i = 1; i <= m; ++i) y[i - 1] = y[i - 1] + temp * AT(a, i, j, lda);
    }
  } else {
    for (int j = 1; j <= n; ++j) {
      double temp = 0.0;
      for (int i = 1; i <= m; ++i) temp = temp + AT(a, i, j, lda) * x[i - 1];
      y[j - 1] = y[j - 1] + alpha * temp;
    }
  }
}

// DASUM, incx = 1: the reference's 6-way unrolled loop adds left to right
// (BLAS/SRC/dasum.f:104-120), i.e. in index order.
double asum(int n, const double* x) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + std::fabs(x[i]);
  return s;
}

// DLACN2 (SRC/dlacn2.f:167-282): Hager/Higham 1-norm estimator by reverse
// communication; isave[0..2] is ISAVE(1..3).
void lacn2(int n, double* v, double* x, int* isgn, double* est, int* kase, int* isave) {
  const int itmax = 5;
  if (*kase == 0) {
    for (int i = 0; i < n; ++i) x[i] = 1.0 / (double)n;
    *kase = 1;
    isave[0] = 1;
    return;
  }
  auto main_loop = [&]() {  // label 50
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    x[isave[1] - 1] = 1.0;
    *kase = 1;
    isave[0] = 3;
  };
  auto final_stage = [&]() {  // label 120
    double altsgn = 1.0;
    for (int i = 1; i <= n; ++i) {
      x[i - 1] = altsgn * (1.0 + (double)(i - 1) / (double)(n - 1));
      altsgn = -altsgn;
    }
    *kase = 1;
    isave[0] = 5;
  };
  switch (isave[0]) {
    case 1: {  // label 20
      if (n == 1) {
        v[0] = x[0];
        *est = std::fabs(v[0]);
        *kase = 0;
        return;
      }
      *est = asum(n, x);
      for (int i = 0; i < n; ++i) {
        x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
        isgn[i] = (int)std::lround(x[i]);
      }
      *kase = 2;
      isave[0] = 2;
      return;
    }
    case 2: {  // label 40
      isave[1] = iamax(n, x, 1);
      isave[2] = 2;
      main_loop();
      return;
    }
    case 3: {  // label 70
      for (int i = 0; i < n; ++i) v[i] = x[i];
      const double estold = *est;
      *est = asum(n, v);
      bool same = true;
      for (int i = 0; i < n; ++i) {
        const double xs = (x[i] >= 0.0) ? 1.0 : -1.0;
        if ((int)std::lround(xs) != isgn[i]) { same = false; break; }
      }
      if (same || *est <= estold) { final_stage(); return; }
      for (int i = 0; i < n; ++i) {
        x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
        isgn[i] = (int)std::lround(x[i]);
      }
      *kase = 2;
      isave[0] = 4;
      return;
    }
    case 4: {  // label 110
      const int jlast = isave[1];
      isave[1] = iamax(n, x, 1);
      if (x[jlast - 1] != std::fabs(x[isave[1] - 1]) && isave[2] < itmax) {
        isave[2] = isave[2] + 1;
        main_loop();
        return;
      }
      final_stage();
      return;
    }
    case 5: {  // label 140
      const double temp = 2.0 * (asum(n, x) / (double)(3 * n));
      if (temp > *est) {
        for (int i = 0; i < n; ++i) v[i] = x[i];
        *est = temp;
      }
      *kase = 0;
      return;
    }
  }
}

// DGERFS (SRC/dgerfs.f:262-432): refinement with ITMAX = 5, componentwise
// backward error BERR and the DLACN2 bound FERR, column by column.
int gerfs(char trans, int n, int nrhs, const double* a, int lda, const double* af, int ldaf, const int* ipiv,
          const double* b, int ldb, double* x, int ldx, double* ferr, double* berr) {
  const bool notran = (trans == 'N' || trans == 'n');
  if (!notran && !(trans == 'T' || trans == 't' || trans == 'C' || trans == 'c')) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max(1, n)) return -5;
  if (ldaf < std::max(1, n)) return -7;
  if (ldb < std::max(1, n)) return -10;
  if (ldx < std::max(1, n)) return -12;
  if (n == 0 || nrhs == 0) {
    for (int j = 0; j < nrhs; ++j) { ferr[j] = 0.0; berr[j] = 0.0; }
    return 0;
  }
  const char transt = notran ? 'T' : 'N';
  const int itmax = 5;
  const int nz = n + 1;
  const double eps = Mach<double>::eps;
  const double safmin = Mach<double>::sfmin;
  const double safe1 = nz * safmin;
  const double safe2 = safe1 / eps;
  std::vector<double> work(3 * (size_t)n);
  std::vector<int> iwork(n);
  double* w1 = work.data();
  double* w2 = work.data() + n;
  double* w3 = work.data() + 2 * (size_t)n;
  for (int j = 1; j <= nrhs; ++j) {
    double* xj = &AT(x, 1, j, ldx);
    const double* bj = &AT(b, 1, j, ldb);
    int count = 1;
    double lstres = 3.0;
    while (true) {
      for (int i = 0; i < n; ++i) w2[i] = bj[i];
      gemv(trans, n, n, -1.0, a, lda, xj, 1.0, w2);
      for (int i = 0; i < n; ++i) w1[i] = std::fabs(bj[i]);
      if (notran) {
        for (int k = 1; k <= n; ++k) {
          const double xk = std::fabs(xj[k - 1]);
          for (int i = 1; i <= n; ++i) w1[i - 1] = w1[i - 1] + std::fabs(AT(a, i, k, lda)) * xk;
        }
      } else {
        for (int k = 1; k <= n; ++k) {
          double sacc = 0.0;
          for (int i = 1; i <= n; ++i) sacc = sacc + std::fabs(AT(a, i, k, lda)) * std::fabs(xj[i - 1]);
          w1[k - 1] = w1[k - 1] + sacc;
        }
      }
      double sb = 0.0;
      for (int i = 0; i < n; ++i) {
        if (w1[i] > safe2) sb = std::max(sb, std::fabs(w2[i]) / w1[i]);
        else sb = std::max(sb, (std::fabs(w2[i]) + safe1) / (w1[i] + safe1));
      }
      berr[j - 1] = sb;
      if (berr[j - 1] > eps && 2.0 * berr[j - 1] <= lstres && count <= itmax) {
        getrs(trans, n, 1, af, ldaf, ipiv, w2, n);
        for (int i = 0; i < n; ++i) xj[i] = xj[i] + 1.0 * w2[i];
        lstres = berr[j - 1];
        count = count + 1;
        continue;
      }
      break;
    }
    for (int i = 0; i < n; ++i) {
      if (w1[i] > safe2) w1[i] = std::fabs(w2[i]) + nz * eps * w1[i];
      else w1[i] = std::fabs(w2[i]) + nz * eps * w1[i] + safe1;
    }
    int kase = 0;
    int isave[3] = {0, 0, 0};
    while (true) {
      lacn2(n, w3, w2, iwork.data(), &ferr[j - 1], &kase, isave);
      if (kase == 0) break;
      if (kase == 1) {
        getrs(transt, n, 1, af, ldaf, ipiv, w2, n);
        for (int i = 0; i < n; ++i) w2[i] = w1[i] * w2[i];
      } else {
        for (int i = 0; i < n; ++i) w2[i] = w1[i] * w2[i];
        getrs(trans, n, 1, af, ldaf, ipiv, w2, n);
      }
    }
    lstres = 0.0;
    for (int i = 0; i < n; ++i) lstres = std::max(lstres, std::fabs(xj[i]));
    if (lstres != 0.0) ferr[j - 1] = ferr[j - 1] / lstres;
  }
  return 0;
}

}  // namespace

extern "C" {

// ---- FP64 path -----------------------------------------------------------
int oracle_idamax(int n, const double* x, int incx) { return iamax(n, x, incx); }
int oracle_isamax(int n, const float* x, int incx) { return iamax(n, x, incx); }
void oracle_dscal(int n, double da, double* x, int incx) { scal(n, da, x, incx); }
void oracle_dlaswp(int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  laswp(n, a, lda, k1, k2, ipiv, incx);
}
void oracle_dtrsm_left(char uplo, char trans, char diag, int m, int n, double alpha, const double* a,
                       int lda, double* b, int ldb) {
  trsm_left(uplo, trans, diag, m, n, alpha, a, lda, b, ldb);
}
void oracle_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* a, int lda,
                  const double* b, int ldb, double beta, double* c, int ldc) {
  gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
int oracle_dgetrf2(int m, int n, double* a, int lda, int* ipiv) { return getrf2(m, n, a, lda, ipiv); }
int oracle_dgetrf(int m, int n, double* a, int lda, int* ipiv, int nb) { return getrf(m, n, a, lda, ipiv, nb); }
int oracle_dgetrs(char trans, int n, int nrhs, const double* a, int lda, const int* ipiv, double* b, int ldb) {
  return getrs(trans, n, nrhs, a, lda, ipiv, b, ldb);
}
double oracle_dlange(char norm, int m, int n, const double* a, int lda) { return lange(norm, m, n, a, lda); }

// ---- FP32 twins (SRC/sgetrf.f, sgetrf2.f, sgetrs.f, slaswp.f, BLAS sgemm/strsm) ----
void oracle_slaswp(int n, float* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  laswp(n, a, lda, k1, k2, ipiv, incx);
}
void oracle_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float* a, int lda,
                  const float* b, int ldb, float beta, float* c, int ldc) {
  gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
void oracle_strsm_left(char uplo, char trans, char diag, int m, int n, float alpha, const float* a,
                       int lda, float* b, int ldb) {
  trsm_left(uplo, trans, diag, m, n, alpha, a, lda, b, ldb);
}
int oracle_sgetrf(int m, int n, float* a, int lda, int* ipiv, int nb) { return getrf(m, n, a, lda, ipiv, nb); }
int oracle_sgetrs(char trans, int n, int nrhs, const float* a, int lda, const int* ipiv, float* b, int ldb) {
  return getrs(trans, n, nrhs, a, lda, ipiv, b, ldb);
}
int oracle_dlag2s(int m, int n, const double* a, int lda, float* sa, int ldsa) { return lag2s(m, n, a, lda, sa, ldsa); }
void oracle_slag2d(int m, int n, const float* sa, int ldsa, double* a, int lda) { lag2d(m, n, sa, ldsa, a, lda); }

// DSGESV: FP32 LU + FP64 iterative refinement, FP64 fallback
// (SRC/dsgesv.f:259-427).  work is n*nrhs doubles, swork n*(n+nrhs) floats.
// Returns INFO; *iter receives ITER.  nb is the ILAENV block size (64).
int oracle_dsgesv(int n, int nrhs, double* a, int lda, int* ipiv, const double* b, int ldb, double* x,
                  int ldx, double* work, float* swork, int* iter, int nb) {
  *iter = 0;
  if (n < 0) return -1;
  if (nrhs < 0) return -2;
  if (lda < std::max(1, n)) return -4;
  if (ldb < std::max(1, n)) return -7;
  if (ldx < std::max(1, n)) return -9;
  if (n == 0) return 0;
  const int itermax = 30;
  const double anrm = lange('I', n, n, a, lda);
  const double eps = Mach<double>::eps;
  const double cte = anrm * eps * std::sqrt((double)n) * 1.0;
  float* sa = swork;
  float* sx = swork + (size_t)n * n;
  int info = 0;
  auto residual = [&]() {  // WORK = B - A*X  (dsgesv.f:328-332)
    for (int j = 1; j <= nrhs; ++j)
      for (int i = 1; i <= n; ++i) AT(work, i, j, n) = AT(b, i, j, ldb);
    gemm('N', 'N', n, nrhs, n, -1.0, a, lda, x, ldx, 1.0, work, n);
  };
  auto converged = [&]() {  // dsgesv.f:337-342
    for (int i = 1; i <= nrhs; ++i) {
      const double xnrm = std::fabs(AT(x, iamax(n, &AT(x, 1, i, ldx), 1), i, ldx));
      const double rnrm = std::fabs(AT(work, iamax(n, &AT(work, 1, i, n), 1), i, n));
      if (rnrm > xnrm * cte) return false;
    }
    return true;
  };
  bool fallback = false;
  if (lag2s(n, nrhs, b, ldb, sx, n) != 0) { *iter = -2; fallback = true; }
  if (!fallback && lag2s(n, n, a, lda, sa, n) != 0) { *iter = -2; fallback = true; }
  if (!fallback && getrf(n, n, sa, n, ipiv, nb) != 0) { *iter = -3; fallback = true; }
  if (!fallback) {
    getrs('N', n, nrhs, sa, n, ipiv, sx, n);
    lag2d(n, nrhs, sx, n, x, ldx);
    residual();
    if (converged()) { *iter = 0; return 0; }
    for (int iiter = 1; iiter <= itermax; ++iiter) {
      if (lag2s(n, nrhs, work, n, sx, n) != 0) { *iter = -2; fallback = true; break; }
      getrs('N', n, nrhs, sa, n, ipiv, sx, n);
      lag2d(n, nrhs, sx, n, work, n);
      for (int i = 1; i <= nrhs; ++i)
        for (int r = 1; r <= n; ++r) AT(x, r, i, ldx) = AT(x, r, i, ldx) + 1.0 * AT(work, r, i, n);
      residual();
      if (converged()) { *iter = iiter; return 0; }
    }
    if (!fallback) *iter = -itermax - 1;
  }
  // FP64 fallback (dsgesv.f:413-427)
  info = getrf(n, n, a, lda, ipiv, nb);
  if (info != 0) return info;
  for (int j = 1; j <= nrhs; ++j)
    for (int i = 1; i <= n; ++i) AT(x, i, j, ldx) = AT(b, i, j, ldb);
  return getrs('N', n, nrhs, a, lda, ipiv, x, ldx);
}

// ---- equilibration (SRC/dgeequ.f:175-296, dgeequb.f:182-313, dlaqge.f:173-226) ----
int oracle_dgeequ(int m, int n, const double* a, int lda, double* r, double* c, double* rowcnd,
                  double* colcnd, double* amax, int radix_round) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  if (m == 0 || n == 0) { *rowcnd = 1; *colcnd = 1; *amax = 0; return 0; }
  const double smlnum = Mach<double>::sfmin;
  const double bignum = 1.0 / smlnum;
  const double radix = 2.0;
  const double logrdx = std::log(radix);
  for (int i = 0; i < m; ++i) r[i] = 0;
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= m; ++i) r[i - 1] = std::max(r[i - 1], std::fabs(AT(a, i, j, lda)));
  if (radix_round)
    for (int i = 0; i < m; ++i)
      if (r[i] > 0) r[i] = std::pow(radix, (int)(std::log(r[i]) / logrdx));
  double rcmin = bignum, rcmax = 0;
  for (int i = 0; i < m; ++i) { rcmax = std::max(rcmax, r[i]); rcmin = std::min(rcmin, r[i]); }
  *amax = rcmax;
  if (rcmin == 0) {
    for (int i = 0; i < m; ++i) if (r[i] == 0) return i + 1;
  } else {
    for (int i = 0; i < m; ++i) r[i] = 1.0 / std::min(std::max(r[i], smlnum), bignum);
    *rowcnd = std::max(rcmin, smlnum) / std::min(rcmax, bignum);
  }
  for (int j = 0; j < n; ++j) c[j] = 0;
  for (int j = 1; j <= n; ++j) {
    for (int i = 1; i <= m; ++i) c[j - 1] = std::max(c[j - 1], std::fabs(AT(a, i, j, lda)) * r[i - 1]);
    if (radix_round && c[j - 1] > 0) c[j - 1] = std::pow(radix, (int)(std::log(c[j - 1]) / logrdx));
  }
  rcmin = bignum; rcmax = 0;
  for (int j = 0; j < n; ++j) { rcmin = std::min(rcmin, c[j]); rcmax = std::max(rcmax, c[j]); }
  if (rcmin == 0) {
    for (int j = 0; j < n; ++j) if (c[j] == 0) return m + j + 1;
  } else {
    for (int j = 0; j < n; ++j) c[j] = 1.0 / std::min(std::max(c[j], smlnum), bignum);
    *colcnd = std::max(rcmin, smlnum) / std::min(rcmax, bignum);
  }
  return 0;
}

// Returns EQUED as a character ('N','R','C','B').
char oracle_dlaqge(int m, int n, double* a, int lda, const double* r, const double* c, double rowcnd,
                   double colcnd, double amax) {
  if (m <= 0 || n <= 0) return 'N';
  const double thresh = 0.1;
  const double small = Mach<double>::sfmin / (Mach<double>::eps * 2.0);  // DLAMCH('P') = eps*base
  const double large = 1.0 / small;
  if (rowcnd >= thresh && amax >= small && amax <= large) {
    if (colcnd >= thresh) return 'N';
    for (int j = 1; j <= n; ++j) {
      const double cj = c[j - 1];
      for (int i = 1; i <= m; ++i) AT(a, i, j, lda) = cj * AT(a, i, j, lda);
    }
    return 'C';
  } else if (colcnd >= thresh) {
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m; ++i) AT(a, i, j, lda) = r[i - 1] * AT(a, i, j, lda);
    return 'R';
  }
  for (int j = 1; j <= n; ++j) {
    const double cj = c[j - 1];
    for (int i = 1; i <= m; ++i) AT(a, i, j, lda) = cj * r[i - 1] * AT(a, i, j, lda);
  }
  return 'B';
}

// ---- test ratios (TESTING/LIN) ----------------------------------------------
// DGET01: ||L*U - P*A||_1 / (n * ||A||_1 * eps) (TESTING/LIN/dget01.f:144-207).
// The product L*U is formed with plain loops, then the row interchanges are
// undone with DLASWP(..., -1) as the reference does.
double oracle_dget01(int m, int n, const double* a, int lda, const double* afac, int ldafac,
                     const int* ipiv) {
  if (m <= 0 || n <= 0) return 0.0;
  const double eps = Mach<double>::eps;
  const double anorm = lange('1', m, n, a, lda);
  const int k = std::min(m, n);
  std::vector<double> lu((size_t)m * n, 0.0);
  // (L*U)(i,j) = sum_{p<=min(i,j,k)} L(i,p)*U(p,j), L unit lower (m x k), U upper (k x n).
  for (int j = 1; j <= n; ++j) {
    for (int p = 1; p <= std::min(j, k); ++p) {
      const double upj = AT(afac, p, j, ldafac);
      lu[(size_t)(p - 1) + (size_t)(j - 1) * m] += upj;  // L(p,p) = 1
      for (int i = p + 1; i <= m; ++i) lu[(size_t)(i - 1) + (size_t)(j - 1) * m] += AT(afac, i, p, ldafac) * upj;
    }
  }
  laswp(n, lu.data(), m, 1, k, ipiv, -1);
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= m; ++i) lu[(size_t)(i - 1) + (size_t)(j - 1) * m] -= AT(a, i, j, lda);
  double resid = lange('1', m, n, lu.data(), m);
  if (anorm <= 0) return resid != 0 ? 1.0 / eps : 0.0;
  return ((resid / (double)n) / anorm) / eps;
}

// DGET02 ('N'/'T'): max_j ||b_j - op(A) x_j||_1 / (||A|| ||x_j||_1 eps), no
// factor n (TESTING/LIN/dget02.f:181-219).  b is not modified here.
double oracle_dget02(char trans, int m, int n, int nrhs, const double* a, int lda, const double* x, int ldx,
                     const double* b, int ldb) {
  if (m <= 0 || n <= 0 || nrhs == 0) return 0.0;
  const bool tr = (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c');
  const int n1 = tr ? n : m, n2 = tr ? m : n;
  const double eps = Mach<double>::eps;
  const double anorm = tr ? lange('I', m, n, a, lda) : lange('1', m, n, a, lda);
  if (anorm <= 0) return 1.0 / eps;
  std::vector<double> r((size_t)n1 * nrhs);
  for (int j = 1; j <= nrhs; ++j)
    for (int i = 1; i <= n1; ++i) r[(size_t)(i - 1) + (size_t)(j - 1) * n1] = AT(b, i, j, ldb);
  gemm(tr ? 'T' : 'N', 'N', n1, nrhs, n2, -1.0, a, lda, x, ldx, 1.0, r.data(), n1);
  double resid = 0;
  for (int j = 1; j <= nrhs; ++j) {
    double bnorm = 0, xnorm = 0;
    for (int i = 1; i <= n1; ++i) bnorm += std::fabs(r[(size_t)(i - 1) + (size_t)(j - 1) * n1]);
    for (int i = 1; i <= n2; ++i) xnorm += std::fabs(AT(x, i, j, ldx));
    if (xnorm <= 0) resid = 1.0 / eps;
    else resid = std::max(resid, ((bnorm / anorm) / xnorm) / eps);
  }
  return resid;
}

// DGET08: the infinity-norm form of DGET02 (TESTING/LIN/dget08.f:178-213).
int oracle_dgerfs(char trans, int n, int nrhs, const double* a, int lda, const double* af, int ldaf,
                  const int* ipiv, const double* b, int ldb, double* x, int ldx, double* ferr, double* berr) {
  return gerfs(trans, n, nrhs, a, lda, af, ldaf, ipiv, b, ldb, x, ldx, ferr, berr);
}

double oracle_dget08(char trans, int m, int n, int nrhs, const double* a, int lda, const double* x, int ldx,
                     const double* b, int ldb) {
  if (m <= 0 || n <= 0 || nrhs == 0) return 0.0;
  const bool tr = (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c');
  const int n1 = tr ? n : m, n2 = tr ? m : n;
  const double eps = Mach<double>::eps;
  const double anorm = lange('I', n1, n2, a, lda);
  if (anorm <= 0) return 1.0 / eps;
  std::vector<double> r((size_t)n1 * nrhs);
  for (int j = 1; j <= nrhs; ++j)
    for (int i = 1; i <= n1; ++i) r[(size_t)(i - 1) + (size_t)(j - 1) * n1] = AT(b, i, j, ldb);
  gemm(tr ? 'T' : 'N', 'N', n1, nrhs, n2, -1.0, a, lda, x, ldx, 1.0, r.data(), n1);
  double resid = 0;
  for (int j = 1; j <= nrhs; ++j) {
    const double bnorm = std::fabs(r[(size_t)(iamax(n1, &r[(size_t)(j - 1) * n1], 1) - 1) + (size_t)(j - 1) * n1]);
    const double xnorm = std::fabs(AT(x, iamax(n2, &AT(x, 1, j, ldx), 1), j, ldx));
    if (xnorm <= 0) resid = 1.0 / eps;
    else resid = std::max(resid, ((bnorm / anorm) / xnorm) / eps);
  }
  return resid;
}

// DGET04: max_j ||x_j - xact_j||_inf / ||xact_j||_inf * RCOND / eps, the
// forward error scaled by the condition estimate (TESTING/LIN/dget04.f:
// 126-175): 1/eps when RCOND < 0 or an exact solution of zero is missed.
double oracle_dget04(int n, int nrhs, const double* x, int ldx, const double* xact, int ldxact, double rcond) {
  if (n <= 0 || nrhs <= 0) return 0.0;
  const double eps = Mach<double>::eps;
  if (rcond < 0) return 1.0 / eps;
  double resid = 0.0;
  for (int j = 1; j <= nrhs; ++j) {
    const int ix = iamax(n, &AT(xact, 1, j, ldxact), 1);
    const double xnorm = std::fabs(AT(xact, ix, j, ldxact));
    double diffnm = 0.0;
    for (int i = 1; i <= n; ++i) diffnm = std::max(diffnm, std::fabs(AT(x, i, j, ldx) - AT(xact, i, j, ldxact)));
    if (xnorm <= 0) {
      if (diffnm > 0) resid = 1.0 / eps;
    } else {
      resid = std::max(resid, (diffnm / xnorm) * rcond);
    }
  }
  if (resid * eps < 1.0) resid = resid / eps;
  return resid;
}

// DGET07: the two error-bound ratios of DGERFS (TESTING/LIN/dget07.f:
// 190-281): reslts[0] = max_j (||x_j - xact_j||_inf / ||x_j||_inf) / FERR_j
// (1/eps where FERR does not bound the error; only when chkferr), and
// reslts[1] = max_j BERR_j / ((n+1) eps + (n+1) unfl / max(min_i (|b| +
// |op(A)||x|)_i, (n+1) unfl)).
void oracle_dget07(char trans, int n, int nrhs, const double* a, int lda, const double* b, int ldb,
                   const double* x, int ldx, const double* xact, int ldxact, const double* ferr, int chkferr,
                   const double* berr, double* reslts) {
  if (n <= 0 || nrhs <= 0) {
    reslts[0] = 0.0;
    reslts[1] = 0.0;
    return;
  }
  const double eps = Mach<double>::eps;
  const double unfl = Mach<double>::sfmin;
  const double ovfl = 1.0 / unfl;
  const bool notran = (trans == 'N' || trans == 'n');
  double errbnd = 0.0;
  if (chkferr) {
    for (int j = 1; j <= nrhs; ++j) {
      const int imax = iamax(n, &AT(x, 1, j, ldx), 1);
      const double xnorm = std::max(std::fabs(AT(x, imax, j, ldx)), unfl);
      double diff = 0.0;
      for (int i = 1; i <= n; ++i) diff = std::max(diff, std::fabs(AT(x, i, j, ldx) - AT(xact, i, j, ldxact)));
      if (!(xnorm > 1.0) && !(diff <= ovfl * xnorm)) {
        errbnd = 1.0 / eps;
        continue;
      }
      if (diff / xnorm <= ferr[j - 1]) errbnd = std::max(errbnd, (diff / xnorm) / ferr[j - 1]);
      else errbnd = 1.0 / eps;
    }
  }
  reslts[0] = errbnd;
  for (int k = 1; k <= nrhs; ++k) {
    double axbi = 0.0;
    for (int i = 1; i <= n; ++i) {
      double tmp = std::fabs(AT(b, i, k, ldb));
      if (notran) {
        for (int j = 1; j <= n; ++j) tmp = tmp + std::fabs(AT(a, i, j, lda)) * std::fabs(AT(x, j, k, ldx));
      } else {
        for (int j = 1; j <= n; ++j) tmp = tmp + std::fabs(AT(a, j, i, lda)) * std::fabs(AT(x, j, k, ldx));
      }
      axbi = (i == 1) ? tmp : std::min(axbi, tmp);
    }
    const double tmp = berr[k - 1] / ((n + 1) * eps + (n + 1) * unfl / std::max(axbi, (n + 1) * unfl));
    reslts[1] = (k == 1) ? tmp : std::max(reslts[1], tmp);
  }
}

}  // extern "C"
