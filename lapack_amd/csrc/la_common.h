// la_common.h -- shared definitions for the gfx950 LU solve library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

typedef double la_d4 __attribute__((ext_vector_type(4)));
typedef float la_f4 __attribute__((ext_vector_type(4)));

// Library error codes (negative, below every LAPACK argument code).
enum {
  LA_ERR_HIP = -1001,       // a HIP runtime call failed
  LA_ERR_NOMEM = -1002,     // device or pinned allocation failed
  LA_ERR_SYNC = -1003,      // a cooperative panel kernel timed out on its grid barrier
  LA_ERR_NODEVICE = -1004,  // no GPU visible
  LA_ERR_LIMIT = -1005,     // a size beyond what this build supports
};

namespace la {

// ---- device helpers -------------------------------------------------------

// Device-coherent (agent scope) relaxed loads/stores: global_load/store with
// sc1, visible across the 8 XCDs' L2s without an L2 write-back.
template <class T> __device__ __forceinline__ void st_dev(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T ld_dev(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until every vector-memory operation this wave issued has completed
// (s_waitcnt vmcnt(0); gfx9 encoding with expcnt/lgkmcnt left at max).
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ---- host-side launchers (defined in the .hip files) ---------------------

// C <- alpha*op(A)*op(B) + beta*C, column-major, FP64 MFMA.
void launch_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, long lda,
                  const double* B, long ldb, double beta, double* C, long ldc, hipStream_t s, int nskip = 0);
// The DGETRF trailing update C <- C - A*B (its own kernel symbols).
// nskip: leave XCDs 0 .. nskip-1 to a concurrent panel kernel.
void launch_dgemm_update(int m, int n, int k, const double* A, long lda, const double* B, long ldb, double* C,
                         long ldc, int nskip, hipStream_t s);
// FP32 update with operands rounded to FP16 (kind 1) or BF16 (kind 2), FP32
// accumulation on the half-precision matrix cores (gemm_h.hip).
void launch_sgemm_update_half(int kind, int m, int n, int k, const float* A, long lda, const float* B, long ldb,
                              float* C, long ldc, int nskip, hipStream_t s);
void launch_sgemm_update(int m, int n, int k, const float* A, long lda, const float* B, long ldb, float* C, long ldc,
                         int nskip, hipStream_t s);
// Same in FP32 (MFMA f32).
void launch_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float* A, long lda,
                  const float* B, long ldb, float beta, float* C, long ldc, hipStream_t s, int nskip = 0);

// out(j,i) = in(i,j): in is rows x cols with ld_in, out cols x rows with ld_out.
// skip (optional device int): nothing is written when *skip != 0.
void launch_transpose_d(int rows, int cols, const double* in, long ld_in, double* out, long ld_out,
                        hipStream_t s, const int* skip = nullptr);
void launch_transpose_s(int rows, int cols, const float* in, long ld_in, float* out, long ld_out,
                        hipStream_t s, const int* skip = nullptr);
// Cast/transposing variants used by the mixed-precision solver.
void launch_transpose_d2s(int rows, int cols, const double* in, long ld_in, float* out, long ld_out,
                          int* overflow, hipStream_t s);

// Row interchanges in the transposed workspace T = A^T (A's rows are T's
// columns): the sequence ipiv[k1..k2) (1-based entries, applied in order)
// is composed into at most swp_max_pairs() (dst, src) position pairs over
// [k1, npos), then applied to T rows [r0, r1) and [r2, r3) in one pass.
int swp_max_pairs();
void launch_swaps_to_pairs(int k1, int k2, int npos, const int* ipiv, int* pairs, int* npairs, int* gmap,
                           hipStream_t s);
void launch_apply_pairs_cols_d(double* T, long ldt, int r0, int r1, int r2, int r3, const int* pairs,
                               const int* npairs, hipStream_t s, int nskip = 0);
void launch_apply_pairs_cols_s(float* T, long ldt, int r0, int r1, int r2, int r3, const int* pairs,
                               const int* npairs, hipStream_t s, int nskip = 0);
// Between two leaves of a panel: leaf [jj, jj+w)'s interchanges applied to
// T rows [j, jj) and [jj+w, jend), then the in-panel DTRSM of T rows
// [jj+w, jend); one workgroup.  w <= 64.
void launch_panel_boundary_d(double* T, long ldt, int j, int jj, int w, int jend, int m, const int* ipiv,
                             hipStream_t s);
void launch_panel_boundary_s(float* T, long ldt, int j, int jj, int w, int jend, int m, const int* ipiv,
                             hipStream_t s);
void launch_apply_pairs_rows_d(double* B, long ldb, int ncols, const int* pairs, const int* npairs, hipStream_t s);
void launch_apply_pairs_rows_s(float* B, long ldb, int ncols, const int* pairs, const int* npairs, hipStream_t s);
// DLASWP on a column-major matrix: perm (nrows ints) <- the permutation of
// the interchange sequence ipiv[k1..k2-1] (1-based entries, incx = +1/-1),
// then rows are gathered through it.
void launch_swaps_to_perm(int nrows, int k1, int k2, const int* ipiv, int incx, int* perm, hipStream_t s,
                          int* gp = nullptr);
size_t perm_pairs_ints(int k1, int k2);  // device scratch of launch_swaps_to_perm's two-launch form
// skip (optional device int): nothing is written when *skip != 0.
// A(r,j) <- A_old(perm[r], j) in place for r < nrows, one pass through LDS;
// false (nothing launched) when nrows exceeds permute_rows_max(sizeof(T)).
int permute_rows_max(int elem_bytes);
bool launch_permute_rows_d(int nrows, int ncols, const int* perm, double* A, long lda, hipStream_t s);
bool launch_permute_rows_s(int nrows, int ncols, const int* perm, float* A, long lda, hipStream_t s);
void launch_gather_rows_d(int nrows, int ncols, const int* perm, const double* src, long lds, double* dst,
                          long ldd, hipStream_t s, const int* skip = nullptr);
void launch_gather_rows_s(int nrows, int ncols, const int* perm, const float* src, long lds, float* dst,
                          long ldd, hipStream_t s, const int* skip = nullptr);

// Batched triangular block inverses along the diagonal of an n x n matrix:
// block d (rows/cols d*nb..) of tri(U) inverted into W(:, d*nb..), nb <= 256.
void launch_trtri_blocks_d(bool upper, bool unit, int n, int nb, const double* U, long ldu, double* W,
                           long ldw, hipStream_t s);
void launch_trtri_blocks_s(bool upper, bool unit, int n, int nb, const float* U, long ldu, float* W,
                           long ldw, hipStream_t s);
// Inverse of the unit upper triangular nb x nb block U (ldu), into W (ldw).
void launch_trtri_uu_d(int nb, const double* U, long ldu, double* W, long ldw, hipStream_t s);
void launch_trtri_uu_s(int nb, const float* U, long ldu, float* W, long ldw, hipStream_t s);
// Inverse of the non-unit upper triangular nb x nb block U, into W.
void launch_trtri_un_d(int nb, const double* U, long ldu, double* W, long ldw, hipStream_t s);
void launch_trtri_un_s(int nb, const float* U, long ldu, float* W, long ldw, hipStream_t s);

// Copy a rows x cols column-major block.
// nskip (copy, interchanges, GEMM, update): no work on XCDs 0 .. nskip-1,
// which a concurrent panel keeps for itself.
// skip (optional device int): nothing is written when *skip != 0.
void launch_copy_d(int rows, int cols, const double* src, long lds, double* dst, long ldd, hipStream_t s,
                   int nskip = 0, const int* skip = nullptr);
void launch_copy_s(int rows, int cols, const float* src, long lds, float* dst, long ldd, hipStream_t s,
                   int nskip = 0, const int* skip = nullptr);
void launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s);
void launch_fill_i32(int* p, size_t n, int v, hipStream_t s);
// x[i] uniform in [-1, 1) from a counter-based hash of (seed, i)
void launch_fill_random(double* x, size_t n, unsigned long long seed, hipStream_t s);
// When a bounded wait gave up (*err != 0): *info <- LA_ERR_SYNC and bit 0 of
// *sticky set (either pointer may be null).
void launch_fold_err(const int* err, int* info, int* sticky, hipStream_t s);

// Panel (leaf) factorization on the transposed workspace: see panel.hip.
// xchg (panel_xchg_bytes() bytes) must be zero at the start of a
// factorization and tagbase must grow by >= 64 from one launch to the next.
int panel_max_width(int h);  // widest leaf the panel kernel takes at height h
size_t panel_xchg_bytes(int max_groups);
int panel_groups(int h);     // workgroups used at height h (0 = one workgroup, LDS only)
// mode of a leaf taller than one workgroup holds:
enum {
  PANEL_PLAIN = 0,  // G register-resident workgroups anywhere on the device
  PANEL_XCD0 = 1,   // the same on XCD 0 only (8*G blocks launched, every eighth takes part)
  PANEL_TALL = 2,   // column by column over HBM, no workgroup waits for another
};
int launch_getf2_panel_d(double* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase,
                         double* xchg, int* err, int mode, hipStream_t s);
int launch_getf2_panel_s(float* T, long ldt, int j0, int w, int m, int* ipiv, int* info, unsigned tagbase,
                         double* xchg, int* err, int mode, hipStream_t s);

// Mixed-precision helpers (mp.hip).
void launch_cast_d2s(int m, int n, const double* A, long lda, float* S, long lds, int* ovf, hipStream_t s);
void launch_cast_s2d(int m, int n, const float* S, long lds, double* A, long lda, hipStream_t s);
void launch_axpy(int m, int n, const double* X, long ldx, double* Y, long ldy, hipStream_t s);
void launch_col_absmax(int m, int n, const double* A, long lda, unsigned long long* out, hipStream_t s);
void launch_row_abssum(int m, int n, const double* A, long lda, double* rowsum, hipStream_t s);
void launch_row_absmax(int m, int n, const double* A, long lda, double* r, hipStream_t s);
void launch_col_absmax_scaled(int m, int n, const double* A, long lda, const double* r, double* c, hipStream_t s);
void launch_scale_rc(int m, int n, double* A, long lda, const double* r, const double* c, int mode, hipStream_t s);
void launch_identity(int n, double* A, long lda, hipStream_t s);  // A <- I
// out[j] for each column: mode 0 max|a| (NaN-propagating), 1 sum|a|, 2 sum (a/scale)^2,
// 3 max|a| over rows 0..j (upper triangle)
void launch_col_norm(int m, int n, const double* A, long lda, int mode, double scale, double* out, hipStream_t s);
void launch_scale_cast_d2s(int m, int n, const double* A, long lda, const double* r, const double* c, float* S,
                           long lds, int* ovf, hipStream_t s);
void launch_scale_s2d(int m, int n, const float* S, long lds, const double* c, double* Y, long ldy, int add,
                      hipStream_t s);
void launch_vdiv(int n, const double* x, const double* nrm, double* y, hipStream_t s);

// Skinny FP64 products (gemv.hip); ws: gemv_ws_doubles(m, n, r) doubles.
size_t gemv_ws_doubles(int m, int n, int r);
void launch_dgemv_n(int m, int n, int r, double alpha, const double* A, long lda, const double* X, long ldx,
                    double beta, double* Y, long ldy, double* ws, hipStream_t s);
void launch_dgemv_t(int m, int k, double alpha, const double* V, long ldv, const double* w, double beta, double* h,
                    double* ws, hipStream_t s);
void launch_dgemv_cols(int m, int n, double alpha, const double* A, long lda, const double* x, double beta,
                       double* y, bool abs, hipStream_t s);
void launch_dgemv_n_abs(int m, int n, double alpha, const double* A, long lda, const double* X, double beta,
                        double* Y, double* ws, hipStream_t s);

// Largest LDS allocation of one workgroup on the current device (bytes).
int lds_limit();

// Batched LU (batched.hip): one persistent launch over matrices of uneven
// order; order[] lists matrix indices by decreasing n, counter is one
// device word the launcher zeroes (with a kernel) before the launch.
void launch_fill_uniform_d(double* a, size_t count, unsigned long long seed, hipStream_t s);
void launch_fill_uniform_s(float* a, size_t count, unsigned long long seed, hipStream_t s);
int batched_getrf_configs();
void batched_getrf_config(int idx, int* nb, int* nt);
size_t batched_getrf_smem(int idx, int nmax, int elem);
// grid: workgroups of the persistent launch (batched_*_grid, computed once
// per plan so that the launches themselves make no runtime queries).
template <class T> int batched_getrf_grid(int cfg, int batch, int nmax);
template <class T> int batched_getrs_grid(int batch, int nmax);
template <class T>
int launch_getrf_batched(int cfg, int grid, int batch, int nmax, const int* nv, T* const* av, const int* ldav,
                         int* const* ipv, int* infov, const int* order, unsigned* counter, hipStream_t s);
template <class T>
int launch_getrs_batched(int grid, int batch, int nmax, const int* nv, const int* nrhsv, T* const* av,
                         const int* ldav, int* const* ipv, T* const* bv, const int* ldbv, const int* order,
                         unsigned* counter, hipStream_t s);
template <class T> int batched_getrs_t_grid(int batch, int nmax);
int launch_adjoint_batched(int grid, int batch, int nmax, const int* nv, const int* nrhsv, double* const* luv,
                           const int* ldluv, int* const* ipv, double* const* xv, const int* ldxv,
                           double* const* xbv, const int* ldxbv, double* const* abv, const int* ldabv,
                           int accumulate, const int* order, unsigned* counter, hipStream_t s);

// Triangular sweeps of DGETRS in one launch each (getrs.hip).  which: 0 = L
// (unit lower), 1 = U, 2 = U^T, 3 = L^T; the right-hand sides Xin (n x nrhs,
// ldxi) are read, the solution written to X (ldx; a different buffer, whose
// previous contents are overwritten from the start).  ctl:
// tri_chain_ctl_words(nrhs) unsigned of scratch (cleared by the launcher);
// err: set to 1 if a wait timed out (X then incomplete).
size_t tri_chain_ctl_words(int nrhs);
template <class T>
void launch_tri_chain(int which, int n, int nrhs, const T* A, long lda, const T* Xin, long ldxi, T* X, long ldx,
                      unsigned* ctl, int* err, hipStream_t s);

// Rank-k outer product C <- alpha*A*B^T + beta*C for k <= rankk_max()
// (adjoint.hip), the Abar = -Bbar X^T step of the solve's adjoint.
int rankk_max();
void launch_rankk_nt_d(int m, int n, int k, double alpha, const double* A, long lda, const double* B, long ldb,
                       double beta, double* C, long ldc, hipStream_t s);

}  // namespace la
