// mp.hip -- element-wise and reduction kernels of the mixed-precision solver:
// DLAG2S / SLAG2D (SRC/dlag2s.f:105-151, SRC/slag2d.f:101-133), the norms the
// convergence test uses (DLANGE 'I', SRC/dlange.f:173-186; the IDAMAX column
// maxima of SRC/dsgesv.f:337-342), DAXPY, and the power-of-two
// equilibration (DGEEQUB, SRC/dgeequb.f:213-311; DLAQGE, SRC/dlaqge.f:183-226).
#include "la_common.h"

namespace la {
namespace {

inline int grid_for(long work) {
  long g = (work + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

// SA = REAL(A); *overflow = 1 when |a| > SLAMCH('O') for any entry.
__global__ __launch_bounds__(256) void cast_d2s_kernel(int m, int n, const double* __restrict__ A, long lda,
                                                       float* __restrict__ S, long lds, int* __restrict__ ovf) {
  const long tot = (long)m * n;
  bool bad = false;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const long i = e % m, j = e / m;
    const double v = A[i + j * lda];
    if (v < -3.4028234663852886e38 || v > 3.4028234663852886e38) bad = true;
    S[i + j * lds] = (float)v;
  }
  if (bad) atomicOr(ovf, 1);
}

__global__ __launch_bounds__(256) void cast_s2d_kernel(int m, int n, const float* __restrict__ S, long lds,
                                                       double* __restrict__ A, long lda) {
  const long tot = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const long i = e % m, j = e / m;
    A[i + j * lda] = (double)S[i + j * lds];
  }
}

// y(:,j) += x(:,j) (alpha = 1, SRC/dsgesv.f:381-383)
__global__ __launch_bounds__(256) void axpy_kernel(int m, int n, const double* __restrict__ X, long ldx,
                                                   double* __restrict__ Y, long ldy) {
  const long tot = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const long i = e % m, j = e / m;
    Y[i + j * ldy] = Y[i + j * ldy] + X[i + j * ldx];
  }
}

// out[j] = |A(IDAMAX(m, A(1,j), 1), j)| as the bit pattern of a non-negative
// double (atomicMax on the bits orders like the values).  IDAMAX's strict '>'
// (BLAS/SRC/idamax.f:112-118) skips a NaN anywhere but in row 1, and a NaN
// in row 1 is the result; a positive NaN's bits exceed every number's, so
// row 1's NaN wins the atomicMax and all other NaNs are dropped here.
__global__ __launch_bounds__(256) void col_absmax_kernel(int m, int n, const double* __restrict__ A, long lda,
                                                         unsigned long long* __restrict__ out) {
  const int j = blockIdx.y;
  unsigned long long mx = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long)gridDim.x * 256) {
    const double a = fabs(A[i + (long)j * lda]);
    const unsigned long long b = (unsigned long long)__double_as_longlong(a);
    if ((a == a || i == 0) && b > mx) mx = b;
  }
  __shared__ unsigned long long s[256];
  s[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o && s[threadIdx.x + o] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(&out[j], s[0]);
}

// The row reductions run one thread per row in j order (the reference's
// order), 64-thread workgroups so that every CU takes rows, and RU columns
// loaded at a time so that RU loads per thread are in flight.
constexpr int RU = 16;

// rowsum[i] = sum_j |A(i,j)| (the WORK vector of DLANGE 'I'), j-order per row.
__global__ __launch_bounds__(64) void row_abssum_kernel(int m, int n, const double* __restrict__ A, long lda,
                                                        double* __restrict__ rowsum) {
  const long i = (long)blockIdx.x * 64 + threadIdx.x;
  if (i >= m) return;
  const double* p = A + i;
  double s = 0.0;
  int j = 0;
  for (; j + RU <= n; j += RU) {
    double v[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) v[u] = p[(long)(j + u) * lda];
#pragma unroll
    for (int u = 0; u < RU; ++u) s = s + fabs(v[u]);
  }
  for (; j < n; ++j) s = s + fabs(p[(long)j * lda]);
  rowsum[i] = s;
}

// Row maxima r(i) = max_j |A(i,j)| (DGEEQU(B) first pass).
__global__ __launch_bounds__(64) void row_absmax_kernel(int m, int n, const double* __restrict__ A, long lda,
                                                        double* __restrict__ r) {
  const long i = (long)blockIdx.x * 64 + threadIdx.x;
  if (i >= m) return;
  const double* p = A + i;
  double mx = 0.0;
  int j = 0;
  for (; j + RU <= n; j += RU) {
    double v[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) v[u] = p[(long)(j + u) * lda];
#pragma unroll
    for (int u = 0; u < RU; ++u) mx = fmax(mx, fabs(v[u]));
  }
  for (; j < n; ++j) mx = fmax(mx, fabs(p[(long)j * lda]));
  r[i] = mx;
}

// c(j) = max_i |A(i,j)| * r(i) (DGEEQU(B) second pass), one workgroup per column.
__global__ __launch_bounds__(256) void col_absmax_scaled_kernel(int m, int n, const double* __restrict__ A, long lda,
                                                                const double* __restrict__ r, double* __restrict__ c) {
  const int j = blockIdx.x;
  // eight rows per thread in flight at a time (a maximum is exact in any
  // order, and fmax drops a NaN whichever order it meets it in)
  double mx = 0.0;
  for (long i0 = threadIdx.x; i0 < m; i0 += 256 * 8) {
    double v[8], w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const long i = i0 + 256 * u;
      v[u] = i < m ? A[i + (long)j * lda] : 0.0;
      w[u] = i < m ? r[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i0 + 256 * u < m) mx = fmax(mx, fabs(v[u]) * w[u]);
  }
  __shared__ double s[256];
  s[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) c[j] = s[0];
}

// Per-column quantities of DLANGE (SRC/dlange.f:143-200), one thread per
// column walking it in row order (the reference's order, so sums are
// bitwise the reference's): mode 0 max |a| (NaN if the column holds one),
// 1 sum |a|, 2 sum (a/scale)^2, 3 max |a| over the upper triangle (DLANTR
// 'M', 'U', 'N', SRC/dlantr.f:170-180).
__global__ __launch_bounds__(256) void col_norm_kernel(int m, int n, const double* __restrict__ A, long lda,
                                                       int mode, double scale, double* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const double* col = A + (long)j * lda;
  double v = 0.0;
  if (mode == 0) {
    for (int i = 0; i < m; ++i) {
      const double t = fabs(col[i]);
      if (v < t || t != t) v = t;
    }
  } else if (mode == 3) {  // the upper triangle: rows 0 .. min(m-1, j)
    const int me = m < j + 1 ? m : j + 1;
    for (int i = 0; i < me; ++i) {
      const double t = fabs(col[i]);
      if (v < t || t != t) v = t;
    }
  } else if (mode == 1) {
    for (int i = 0; i < m; ++i) v = v + fabs(col[i]);
  } else {
    for (int i = 0; i < m; ++i) {
      const double t = col[i] / scale;
      v = v + t * t;
    }
  }
  out[j] = v;
}

// A <- I (n x n).
__global__ __launch_bounds__(256) void identity_kernel(int n, double* __restrict__ A, long lda) {
  const long tot = (long)n * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const long i = e % n, j = e / n;
    A[i + j * lda] = i == j ? 1.0 : 0.0;
  }
}

// A(i,j) <- c(j)*r(i)*A(i,j) (or one factor): the scaling forms of DLAQGE.
__global__ __launch_bounds__(256) void scale_rc_kernel(int m, int n, double* __restrict__ A, long lda,
                                                       const double* __restrict__ r, const double* __restrict__ c,
                                                       int mode) {
  const long tot = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const long i = e % m, j = e / m;
    double& a = A[i + j * lda];
    if (mode == 1) a = r[i] * a;                 // EQUED = 'R'
    else if (mode == 2) a = c[j] * a;            // 'C'
    else if (mode == 3) a = c[j] * r[i] * a;     // 'B'
  }
}

// S(i,j) = REAL(r(i) * c(j) * A(i,j)) (r, c optional, powers of two in the
// equilibrated solver, so the scaling itself is exact); overflow flag as DLAG2S.
__global__ __launch_bounds__(256) void scale_cast_d2s_kernel(int m, int n, const double* __restrict__ A, long lda,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ c, float* __restrict__ S,
                                                             long lds, int* __restrict__ ovf) {
  const long tot = (long)m * n;
  bool bad = false;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const long i = e % m, j = e / m;
    double v = A[i + j * lda];
    if (r) v = r[i] * v;
    if (c) v = c[j] * v;
    if (v < -3.4028234663852886e38 || v > 3.4028234663852886e38) bad = true;
    S[i + j * lds] = (float)v;
  }
  if (bad && ovf) atomicOr(ovf, 1);
}

// Y(i,j) = c(i) * S(i,j)  (+ Y(i,j) when add): back to FP64, column scaling
// of the equilibrated system applied to the solution (c optional).
__global__ __launch_bounds__(256) void scale_s2d_kernel(int m, int n, const float* __restrict__ S, long lds,
                                                        const double* __restrict__ c, double* __restrict__ Y,
                                                        long ldy, int add) {
  const long tot = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const long i = e % m, j = e / m;
    double v = (double)S[i + j * lds];
    if (c) v = c[i] * v;
    double* yp = Y + i + j * ldy;
    *yp = add ? *yp + v : v;
  }
}

// y = x / *nrm (GMRES basis vector); nrm is a device scalar.
__global__ __launch_bounds__(256) void vdiv_kernel(int n, const double* __restrict__ x,
                                                   const double* __restrict__ nrm, double* __restrict__ y) {
  const double d = *nrm;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) y[i] = x[i] / d;
}

}  // namespace

void launch_scale_cast_d2s(int m, int n, const double* A, long lda, const double* r, const double* c, float* S,
                           long lds, int* ovf, hipStream_t s) {
  if (m > 0 && n > 0) scale_cast_d2s_kernel<<<grid_for((long)m * n), 256, 0, s>>>(m, n, A, lda, r, c, S, lds, ovf);
}
void launch_scale_s2d(int m, int n, const float* S, long lds, const double* c, double* Y, long ldy, int add,
                      hipStream_t s) {
  if (m > 0 && n > 0) scale_s2d_kernel<<<grid_for((long)m * n), 256, 0, s>>>(m, n, S, lds, c, Y, ldy, add);
}
void launch_vdiv(int n, const double* x, const double* nrm, double* y, hipStream_t s) {
  if (n > 0) vdiv_kernel<<<grid_for(n), 256, 0, s>>>(n, x, nrm, y);
}

void launch_cast_d2s(int m, int n, const double* A, long lda, float* S, long lds, int* ovf, hipStream_t s) {
  if (m > 0 && n > 0) cast_d2s_kernel<<<grid_for((long)m * n), 256, 0, s>>>(m, n, A, lda, S, lds, ovf);
}
void launch_cast_s2d(int m, int n, const float* S, long lds, double* A, long lda, hipStream_t s) {
  if (m > 0 && n > 0) cast_s2d_kernel<<<grid_for((long)m * n), 256, 0, s>>>(m, n, S, lds, A, lda);
}
void launch_axpy(int m, int n, const double* X, long ldx, double* Y, long ldy, hipStream_t s) {
  if (m > 0 && n > 0) axpy_kernel<<<grid_for((long)m * n), 256, 0, s>>>(m, n, X, ldx, Y, ldy);
}
void launch_col_absmax(int m, int n, const double* A, long lda, unsigned long long* out, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  launch_fill_i32(reinterpret_cast<int*>(out), (size_t)2 * n, 0, s);  // a kernel, not a memset (graph-safe)
  int gx = (m + 255) / 256;
  if (gx > 64) gx = 64;
  col_absmax_kernel<<<dim3(gx, n), 256, 0, s>>>(m, n, A, lda, out);
}
void launch_row_abssum(int m, int n, const double* A, long lda, double* rowsum, hipStream_t s) {
  if (m > 0) row_abssum_kernel<<<(m + 63) / 64, 64, 0, s>>>(m, n, A, lda, rowsum);
}
void launch_row_absmax(int m, int n, const double* A, long lda, double* r, hipStream_t s) {
  if (m > 0) row_absmax_kernel<<<(m + 63) / 64, 64, 0, s>>>(m, n, A, lda, r);
}
void launch_col_absmax_scaled(int m, int n, const double* A, long lda, const double* r, double* c, hipStream_t s) {
  if (n > 0) col_absmax_scaled_kernel<<<n, 256, 0, s>>>(m, n, A, lda, r, c);
}
void launch_scale_rc(int m, int n, double* A, long lda, const double* r, const double* c, int mode, hipStream_t s) {
  if (m > 0 && n > 0 && mode > 0) scale_rc_kernel<<<grid_for((long)m * n), 256, 0, s>>>(m, n, A, lda, r, c, mode);
}
void launch_col_norm(int m, int n, const double* A, long lda, int mode, double scale, double* out, hipStream_t s) {
  if (n > 0) col_norm_kernel<<<(n + 255) / 256, 256, 0, s>>>(m, n, A, lda, mode, scale, out);
}
void launch_identity(int n, double* A, long lda, hipStream_t s) {
  if (n > 0) identity_kernel<<<grid_for((long)n * n), 256, 0, s>>>(n, A, lda);
}

}  // namespace la
