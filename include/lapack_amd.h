/* lapack_amd.h -- C ABI of the MI355X (gfx950) LU solve library.
 *
 * Three layers, all plain C types (no HIP or torch types in signatures;
 * streams are passed as void* holding a hipStream_t, NULL = the library's
 * own stream for the device current on the calling thread):
 *
 *  1. Fortran LAPACK drop-ins: the exact symbols and calling convention of
 *     the reference (all arguments by reference, CHARACTER lengths appended
 *     as size_t), prototypes as in LAPACKE/include/lapack.h.  Column-major,
 *     in place, IPIV a 1-based interchange sequence.  Pointers may be host
 *     or device memory (detected per pointer); every call returns with the
 *     results in place.  Illegal arguments call XERBLA(name, -INFO) exactly
 *     as the reference does; xerbla_ here is a weak symbol, so a program's
 *     own XERBLA replaces it at link time (TESTING/LIN/xerbla.f:74-118).
 *
 *  2. Value-argument C entry points lapack_amd_*: same semantics, arguments
 *     by value, INFO returned (negative = illegal argument number, positive =
 *     first exact zero pivot, below -1000 = a library/HIP failure, see
 *     LA_ERR_*).  No XERBLA call.
 *
 *  3. Device entry points lapack_amd_*_async: device pointers only, enqueued
 *     on the given stream and returning at once; INFO is written to a device
 *     int.  The headline measurements use these (inputs resident in HBM).
 */
#ifndef LAPACK_AMD_H
#define LAPACK_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library error codes (INFO values below every LAPACK argument code) -- */
#define LA_ERR_HIP (-1001)      /* a HIP runtime call failed                   */
#define LA_ERR_NOMEM (-1002)    /* device allocation failed                    */
#define LA_ERR_SYNC (-1003)     /* a bounded wait between workgroups gave up   */
#define LA_ERR_NODEVICE (-1004) /* no GPU visible                              */
#define LA_ERR_LIMIT (-1005)    /* a size beyond what this build supports      */

/* ======================= 1. Fortran LAPACK drop-ins ======================= */

/* DGETRF: A = P*L*U, blocked right-looking with partial pivoting.
 * Replaces SRC/dgetrf.f:105 (prototype LAPACKE/include/lapack.h:4349-4353). */
void dgetrf_(const int* m, const int* n, double* a, const int* lda, int* ipiv, int* info);

/* DGETRF2: the recursive panel routine, same contract as DGETRF.
 * Replaces SRC/dgetrf2.f:112 (LAPACKE/include/lapack.h:4373). */
void dgetrf2_(const int* m, const int* n, double* a, const int* lda, int* ipiv, int* info);

/* DGETRS: solve A*X = B or A**T*X = B with the factors of DGETRF.
 * Replaces SRC/dgetrs.f:118 (LAPACKE/include/lapack.h:4436-4450). */
void dgetrs_(const char* trans, const int* n, const int* nrhs, const double* a, const int* lda,
             const int* ipiv, double* b, const int* ldb, int* info, size_t trans_len);

/* DGESV: DGETRF then DGETRS('N').  Replaces SRC/dgesv.f:148. */
void dgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b,
            const int* ldb, int* info);

/* DLASWP: row interchanges.  Replaces SRC/dlaswp.f:112 (LAPACKE/include/lapack.h:11326-11328). */
void dlaswp_(const int* n, double* a, const int* lda, const int* k1, const int* k2, const int* ipiv,
             const int* incx);

/* SGETRF / SGETRS: the single-precision twins (SRC/sgetrf.f, SRC/sgetrs.f). */
void sgetrf_(const int* m, const int* n, float* a, const int* lda, int* ipiv, int* info);
void sgetrs_(const char* trans, const int* n, const int* nrhs, const float* a, const int* lda,
             const int* ipiv, float* b, const int* ldb, int* info, size_t trans_len);

/* DSGESV: FP32 LU + FP64 iterative refinement, FP64 fallback.
 * Replaces SRC/dsgesv.f:191 (LAPACKE/include/lapack.h:3727-3735).  WORK
 * (N*NRHS doubles) and SWORK (N*(N+NRHS) floats) are accepted for ABI
 * compatibility; the GPU solver keeps its own device workspace. */
void dsgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, const double* b,
             const int* ldb, double* x, const int* ldx, double* work, float* swork, int* iter,
             int* info);

/* DGERFS: iterative refinement of a solution with forward/backward error
 * bounds, from the factors of DGETRF.  Replaces SRC/dgerfs.f:185.  WORK and
 * IWORK are accepted for ABI compatibility; FERR/BERR may be host or device. */
void dgerfs_(const char* trans, const int* n, const int* nrhs, const double* a, const int* lda, const double* af,
             const int* ldaf, const int* ipiv, const double* b, const int* ldb, double* x, const int* ldx,
             double* ferr, double* berr, double* work, int* iwork, int* info, size_t trans_len);

/* DGEEQU / DGEEQUB: row and column scale factors (DGEEQUB: powers of the
 * radix, so scaling is exact).  Replace SRC/dgeequ.f:137 and
 * SRC/dgeequb.f:144 (LAPACKE/include/lapack.h, dgeequ_/dgeequb_).  R, C may
 * be host or device arrays; the maxima are taken on the GPU. */
void dgeequ_(const int* m, const int* n, const double* a, const int* lda, double* r, double* c, double* rowcnd,
             double* colcnd, double* amax, int* info);
void dgeequb_(const int* m, const int* n, const double* a, const int* lda, double* r, double* c, double* rowcnd,
              double* colcnd, double* amax, int* info);
/* DLAQGE: apply the scaling when ROWCND/COLCND/AMAX call for it; EQUED out.
 * Replaces SRC/dlaqge.f:140. */
void dlaqge_(const int* m, const int* n, double* a, const int* lda, const double* r, const double* c,
             const double* rowcnd, const double* colcnd, const double* amax, char* equed, size_t equed_len);

/* DLANGE: 'M' max |a|, '1'/'O' one norm, 'I' infinity norm, 'F'/'E'
 * Frobenius norm of a host or device matrix, computed on the GPU (sums in
 * the reference's order).  Replaces SRC/dlange.f:114. */
double dlange_(const char* norm, const int* m, const int* n, const double* a, const int* lda, double* work,
               size_t norm_len);

/* DGECON: reciprocal condition number estimate from the factors of DGETRF
 * (DLACN2 on the host, the triangular solves on the GPU).  Replaces
 * SRC/dgecon.f:130 (LAPACKE/include/lapack.h, dgecon_). */
void dgecon_(const char* norm, const int* n, const double* a, const int* lda, const double* anorm, double* rcond,
             double* work, int* iwork, int* info, size_t norm_len);

/* DGETRI: inverse from the factors of DGETRF.  Replaces SRC/dgetri.f:114
 * (LAPACKE/include/lapack.h, dgetri_); WORK(1) returns N*64 as the
 * reference's optimal size, LWORK = -1 is a size query. */
void dgetri_(const int* n, double* a, const int* lda, const int* ipiv, double* work, const int* lwork, int* info);

/* DGESVX: the expert driver -- equilibration, LU, RCOND, solve, refinement
 * with error bounds.  Replaces SRC/dgesvx.f:349 (LAPACKE/include/lapack.h,
 * dgesvx_).  WORK(1) returns the reciprocal pivot growth. */
void dgesvx_(const char* fact, const char* trans, const int* n, const int* nrhs, double* a, const int* lda, double* af,
             const int* ldaf, int* ipiv, char* equed, double* r, double* c, double* b, const int* ldb, double* x,
             const int* ldx, double* rcond, double* ferr, double* berr, double* work, int* iwork, int* info,
             size_t fact_len, size_t trans_len, size_t equed_len);

/* DGEMM: C <- alpha*op(A)*op(B) + beta*C on the FP64 matrix cores.
 * Replaces BLAS/SRC/dgemm.f:213. */
void dgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k,
            const double* alpha, const double* a, const int* lda, const double* b, const int* ldb,
            const double* beta, double* c, const int* ldc, size_t transa_len, size_t transb_len);

/* XERBLA: weak default that reproduces SRC/xerbla.f:86-88 (message, then
 * STOP) unless recording mode is on (lapack_amd_xerbla_record). */
void xerbla_(const char* srname, const int* info, size_t srname_len);

/* ===================== 2. value-argument C entry points =================== */

int lapack_amd_dgetrf(int m, int n, double* a, int lda, int* ipiv);
int lapack_amd_dgetrs(char trans, int n, int nrhs, const double* a, int lda, const int* ipiv, double* b,
                      int ldb);
int lapack_amd_dgesv(int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb);
int lapack_amd_dlaswp(int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx);
int lapack_amd_sgetrf(int m, int n, float* a, int lda, int* ipiv);
int lapack_amd_sgetrs(char trans, int n, int nrhs, const float* a, int lda, const int* ipiv, float* b,
                      int ldb);
int lapack_amd_dgemm(char transa, char transb, int m, int n, int k, double alpha, const double* a, int lda,
                     const double* b, int ldb, double beta, double* c, int ldc);
/* DSGESV semantics (SRC/dsgesv.f:259-427); *iter receives ITER. */
int lapack_amd_dsgesv(int n, int nrhs, double* a, int lda, int* ipiv, const double* b, int ldb, double* x,
                      int ldx, int* iter);

/* Mixed-precision solve to FP64 accuracy.  No single reference routine: it
 * generalizes DSGESV (SRC/dsgesv.f:259-427; same arguments, same per-column
 * stopping test max|r| <= max|x|*||A||_inf*eps*sqrt(n), same FP64 LU fallback
 * and ITER codes -2/-3/-31) to a factorization in PREC:
 *   'S' FP32 LU;  'H' FP32-stored LU whose trailing updates take FP16
 *   operands, 'B' BF16 operands (FP32 accumulation on the matrix cores),
 *   both after DGEEQUB power-of-two equilibration.
 * Refinement in FP64: classic iterations (corrections solved with the
 * low-precision factors in FP32) while the backward error drops tenfold per
 * step, then restarted GMRES on the correction equation preconditioned by
 * the low-precision LU applied in FP64 (GMRES-IR).  *iter = refinement steps + GMRES cycles, *gmres_iter =
 * GMRES inner iterations, berr[j] (optional) = max|r_j|/(||A||_inf max|x_j|),
 * NaN when the FP64 fallback ran (A then holds its factors, as in DSGESV).
 * Returns INFO (PREC is parameter 10). */
int lapack_amd_dgesv_ir(int n, int nrhs, double* a, int lda, int* ipiv, const double* b, int ldb, double* x,
                        int ldx, char prec, int* iter, int* gmres_iter, double* berr);
int lapack_amd_dgerfs(char trans, int n, int nrhs, const double* a, int lda, const double* af, int ldaf,
                      const int* ipiv, const double* b, int ldb, double* x, int ldx, double* ferr, double* berr);
int lapack_amd_dgeequ(int m, int n, const double* a, int lda, double* r, double* c, double* rowcnd, double* colcnd,
                      double* amax);
double lapack_amd_dlange(char norm, int m, int n, const double* a, int lda);
int lapack_amd_dgecon(char norm, int n, const double* a, int lda, double anorm, double* rcond);
int lapack_amd_dgetri(int n, double* a, int lda, const int* ipiv);
/* DGESVX semantics; *rpvgrw receives WORK(1), the reciprocal pivot growth. */
int lapack_amd_dgesvx(char fact, char trans, int n, int nrhs, double* a, int lda, double* af, int ldaf, int* ipiv,
                      char* equed, double* r, double* c, double* b, int ldb, double* x, int ldx, double* rcond,
                      double* ferr, double* berr, double* rpvgrw);
int lapack_amd_dgeequb(int m, int n, const double* a, int lda, double* r, double* c, double* rowcnd,
                       double* colcnd, double* amax);
char lapack_amd_dlaqge(int m, int n, double* a, int lda, const double* r, const double* c, double rowcnd,
                       double colcnd, double amax);

/* ========================= 3. device entry points ========================= */

int lapack_amd_dgetrf_async(int m, int n, double* d_a, int lda, int* d_ipiv, int* d_info, void* stream);
int lapack_amd_dgetrs_async(char trans, int n, int nrhs, const double* d_a, int lda, const int* d_ipiv,
                            double* d_b, int ldb, void* stream);
/* No counterpart in the reference: DGETRS without any device allocation, for
 * callers that cannot allocate (code under HIP graph capture, serving loops
 * or harnesses that forbid hipMalloc).  Same arguments, checks and kernels as
 * lapack_amd_dgetrs_async, plus d_info (a device int): 0, or LA_ERR_SYNC when
 * a bounded wait gave up (B then left as passed).  The workspace is a static
 * device array of lapack_amd_noalloc_scratch_bytes() (16 MiB) per GPU; a call
 * needing more than that (lapack_amd_dgetrs_noalloc_bytes(n, nrhs): n ints,
 * 2 n * nrhs doubles and a few KB, about 4.3 MB at n = 16384, nrhs = 16;
 * n * nrhs up to about 1e6) returns LA_ERR_LIMIT.  Calls that overlap on
 * one GPU share the array and race; order them on one stream. */
int lapack_amd_dgetrs_noalloc_async(char trans, int n, int nrhs, const double* d_a, int lda, const int* d_ipiv,
                                    double* d_b, int ldb, int* d_info, void* stream);
size_t lapack_amd_dgetrs_noalloc_bytes(int n, int nrhs);
size_t lapack_amd_noalloc_scratch_bytes(void);
int lapack_amd_sgetrf_async(int m, int n, float* d_a, int lda, int* d_ipiv, int* d_info, void* stream);
int lapack_amd_sgetrs_async(char trans, int n, int nrhs, const float* d_a, int lda, const int* d_ipiv,
                            float* d_b, int ldb, void* stream);
int lapack_amd_dlaswp_async(int n, double* d_a, int lda, int k1, int k2, const int* d_ipiv, int incx,
                            void* stream);
int lapack_amd_dgemm_async(char transa, char transb, int m, int n, int k, double alpha, const double* d_a,
                           int lda, const double* d_b, int ldb, double beta, double* d_c, int ldc, void* stream);
int lapack_amd_sgemm_async(char transa, char transb, int m, int n, int k, float alpha, const float* d_a,
                           int lda, const float* d_b, int ldb, float beta, float* d_c, int ldc, void* stream);

/* Adjoint of the solve X = A^-1 B (reverse-mode derivative), from the saved
 * DGETRF factors d_lu/d_ipiv and the solution d_x, for the loss gradient
 * d_xbar:  Bbar = A^-T Xbar  -- DGETRS('T') on the saved factors, no
 * refactorization (SRC/dgetrs.f:199-217) --  and Abar = -Bbar X^T, a
 * rank-NRHS DGEMM('N','T') (BLAS/SRC/dgemm.f), added to d_abar when
 * accumulate != 0.  d_bbar may equal d_xbar (in place).  The reference has
 * no single routine for it; it is the composition of the two calls above.
 * Returns 0 or -(parameter number) (N 1, NRHS 2, LDA 4, LDX 7, LDXBAR 9,
 * LDBBAR 11, LDABAR 13). */
int lapack_amd_dgesv_adjoint_async(int n, int nrhs, const double* d_lu, int lda, const int* d_ipiv,
                                   const double* d_x, int ldx, const double* d_xbar, int ldxbar, double* d_bbar,
                                   int ldbbar, double* d_abar, int ldabar, int accumulate, void* stream);

/* ============================ 4. batched LU =============================== */

/* Thousands of independent systems of uneven order n_i (any mix of orders
 * up to the LDS limit: n <= 2500 in FP64, 5000 in FP32), each a
 * column-major matrix of its own with its own LDA, addressed through device
 * arrays of pointers.  The reference has no batched routine; each matrix
 * gets exactly DGETRF (SRC/dgetrf.f:105) / DGETRS('N') (SRC/dgetrs.f:118)
 * semantics: IPIV 1-based per matrix, INFO per matrix.
 *
 * A plan fixes the orders (host array n[batch]) and the precision ('d' or
 * 's'), orders the matrices by decreasing n for the persistent launch, and
 * -- with autotune != 0 -- times the kernel configurations (panel width NB,
 * workgroup size) on synthetic matrices of the plan's sizes and keeps the
 * fastest.  Creating a plan allocates; the *_batched_async launches do
 * not, and may be captured in a hipGraph and replayed. */
typedef struct lapack_amd_batch_plan lapack_amd_batch_plan;
int lapack_amd_batch_plan_create(lapack_amd_batch_plan** plan, int batch, const int* n, char precision,
                                 int autotune);
void lapack_amd_batch_plan_destroy(lapack_amd_batch_plan* plan);
int lapack_amd_batch_plan_config(const lapack_amd_batch_plan* plan, int* nb, int* nthreads, double* tuned_ms);
int lapack_amd_batch_plan_set_config(lapack_amd_batch_plan* plan, int nb, int nthreads);
int lapack_amd_dgetrf_batched_async(const lapack_amd_batch_plan* plan, double* const* d_a, const int* d_lda,
                                    int* const* d_ipiv, int* d_info, void* stream);
int lapack_amd_sgetrf_batched_async(const lapack_amd_batch_plan* plan, float* const* d_a, const int* d_lda,
                                    int* const* d_ipiv, int* d_info, void* stream);
int lapack_amd_dgetrs_batched_async(const lapack_amd_batch_plan* plan, const int* d_nrhs, double* const* d_a,
                                    const int* d_lda, int* const* d_ipiv, double* const* d_b, const int* d_ldb,
                                    void* stream);
int lapack_amd_sgetrs_batched_async(const lapack_amd_batch_plan* plan, const int* d_nrhs, float* const* d_a,
                                    const int* d_lda, int* const* d_ipiv, float* const* d_b, const int* d_ldb,
                                    void* stream);

/* Batched adjoint of the solve (the reverse pass of X_b = A_b^-1 B_b for every
 * system of a 'd' plan), from the saved factors of
 * lapack_amd_dgetrf_batched_async: Xbar_b is overwritten by
 * Bbar_b = A_b^-T Xbar_b -- DGETRS('T') semantics on the saved LU and IPIV,
 * SRC/dgetrs.f:199-217, no refactorization -- and Abar_b <- -Bbar_b X_b^T
 * (+ Abar_b when accumulate != 0), the rank-NRHS DGEMM('N','T') of
 * BLAS/SRC/dgemm.f.  The reference has no batched routine.  Two launches
 * (a persistent transposed solve, one workgroup per system, and the
 * rank-NRHS product over every system), no allocation, capturable in a
 * hipGraph.  Returns 0, -1 for a plan that is not FP64, or LA_ERR_LIMIT. */
int lapack_amd_dgesv_adjoint_batched_async(const lapack_amd_batch_plan* plan, const int* d_nrhs, double* const* d_lu,
                                           const int* d_lda, int* const* d_ipiv, double* const* d_x, const int* d_ldx,
                                           double* const* d_xbar, const int* d_ldxbar, double* const* d_abar,
                                           const int* d_ldabar, int accumulate, void* stream);

/* ====================== 5. LAPACKE C interface ============================ */

/* The LAPACKE functions of the path (LAPACKE/include/lapacke.h; lapack_int
 * is int, the LP64 build): the layout check (-1), the NaN scan of the
 * inputs (LAPACKE_set_nancheck / the LAPACKE_NANCHECK environment variable),
 * INFO of the Fortran routine shifted by one when negative, and for
 * LAPACK_ROW_MAJOR the leading-dimension checks and column-major
 * temporaries -- as LAPACKE/src/lapacke_dgetrf.c, lapacke_dgetrf_work.c and
 * their siblings.  Host or device arrays (row-major device arrays are
 * transposed on the GPU). */
#ifndef LAPACK_ROW_MAJOR
#define LAPACK_ROW_MAJOR 101
#define LAPACK_COL_MAJOR 102
#endif
#ifndef LAPACK_WORK_MEMORY_ERROR
#define LAPACK_WORK_MEMORY_ERROR (-1010)
#define LAPACK_TRANSPOSE_MEMORY_ERROR (-1011)
#endif
void LAPACKE_xerbla(const char* name, int info);
void LAPACKE_set_nancheck(int flag);
int LAPACKE_get_nancheck(void);
int LAPACKE_dgetrf(int matrix_layout, int m, int n, double* a, int lda, int* ipiv);
int LAPACKE_dgetrf_work(int matrix_layout, int m, int n, double* a, int lda, int* ipiv);
int LAPACKE_dgetrf2(int matrix_layout, int m, int n, double* a, int lda, int* ipiv);
int LAPACKE_dgetrf2_work(int matrix_layout, int m, int n, double* a, int lda, int* ipiv);
int LAPACKE_dgetrs(int matrix_layout, char trans, int n, int nrhs, const double* a, int lda, const int* ipiv,
                   double* b, int ldb);
int LAPACKE_dgetrs_work(int matrix_layout, char trans, int n, int nrhs, const double* a, int lda, const int* ipiv,
                        double* b, int ldb);
int LAPACKE_dgesv(int matrix_layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb);
int LAPACKE_dgesv_work(int matrix_layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb);
int LAPACKE_dgetri(int matrix_layout, int n, double* a, int lda, const int* ipiv);
int LAPACKE_dgetri_work(int matrix_layout, int n, double* a, int lda, const int* ipiv, double* work, int lwork);
int LAPACKE_dsgesv(int matrix_layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb, double* x,
                   int ldx, int* iter);
int LAPACKE_dsgesv_work(int matrix_layout, int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb,
                        double* x, int ldx, double* work, float* swork, int* iter);
int LAPACKE_dgecon(int matrix_layout, char norm, int n, const double* a, int lda, double anorm, double* rcond);
int LAPACKE_dgecon_work(int matrix_layout, char norm, int n, const double* a, int lda, double anorm, double* rcond,
                        double* work, int* iwork);
double LAPACKE_dlange(int matrix_layout, char norm, int m, int n, const double* a, int lda);
double LAPACKE_dlange_work(int matrix_layout, char norm, int m, int n, const double* a, int lda, double* work);
int LAPACKE_dlaswp(int matrix_layout, int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx);
int LAPACKE_dlaswp_work(int matrix_layout, int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx);
int LAPACKE_dgesvx(int matrix_layout, char fact, char trans, int n, int nrhs, double* a, int lda, double* af,
                   int ldaf, int* ipiv, char* equed, double* r, double* c, double* b, int ldb, double* x, int ldx,
                   double* rcond, double* ferr, double* berr, double* rpivot);
int LAPACKE_dgesvx_work(int matrix_layout, char fact, char trans, int n, int nrhs, double* a, int lda, double* af,
                        int ldaf, int* ipiv, char* equed, double* r, double* c, double* b, int ldb, double* x,
                        int ldx, double* rcond, double* ferr, double* berr, double* work, int* iwork);

/* ============================ tuning and tools ============================ */

/* Block size NB of the outer loop (the ILAENV(1,'DGETRF') value, reference
 * default 64 at SRC/ilaenv.f:289-295; this build's default 256) and panel
 * leaf width (<= 64).  0 restores the default.  The analogue of the test
 * suite's XLAENV hook (TESTING/LIN/xlaenv.f:77). */
void lapack_amd_set_tuning(int nb, int leaf);
void lapack_amd_get_tuning(int* nb, int* leaf);

/* Autotuner of the plain factorization (the role ILAENV's NB,
 * SRC/ilaenv.f:289-295, plays in the reference): times DGETRF of an n x n
 * matrix on the current device for each candidate NB in {256, 192, 128} and
 * leaf width in {64, 32}, keeps the fastest for n's bucket (the next power
 * of two >= n), and with a non-null path writes the table there as JSON
 * {"getrf": {"<bucket>": {"nb": .., "leaf": .., "ms": ..}}} (merged with the
 * buckets the file already holds).  DGETRF then uses the table's entry for
 * min(m, n)'s bucket unless lapack_amd_set_tuning or LAPACK_AMD_NB /
 * LAPACK_AMD_LEAF say otherwise.  Allocates two n x n matrices while it
 * runs.  Returns 0 or a negative LA_ERR_* code. */
int lapack_amd_tune_getrf(int n, const char* path, int* nb, int* leaf, double* ms);
/* Reads a table written by lapack_amd_tune_getrf (the file named by the
 * environment variable LAPACK_AMD_TUNE_FILE is read at the first
 * factorization); returns the number of buckets read, -1 if unreadable.
 * A null path empties the table. */
int lapack_amd_load_tuning(const char* path);
/* NB and leaf width a factorization with min(m, n) = n would use. */
void lapack_amd_get_tuning_for(int n, int* nb, int* leaf);

/* XERBLA recording mode (the role of TESTING/LIN/xerbla.f): when on, the
 * default xerbla_ records the routine name and parameter number and
 * returns instead of stopping. */
void lapack_amd_xerbla_record(int on);
int lapack_amd_xerbla_last(char* name_out, int cap);
int lapack_amd_xerbla_calls(void);

/* Per-phase GPU time of getrf/getrs, accumulated over calls while enabled
 * (hipEvents on the stream each phase runs on).  lapack_amd_phase_enable:
 * 0 off, 1 every phase, 2 only the trailing update (LA_PHASE_GEMM: two
 * events per update launch, little cost for a timed run). */
enum {
  LA_PHASE_TRANSPOSE = 0,
  LA_PHASE_PANEL = 1,
  LA_PHASE_LASWP = 2,
  LA_PHASE_TRSM = 3,
  LA_PHASE_GEMM_PANEL = 4,
  LA_PHASE_GEMM = 5,
  LA_PHASE_GETRS = 6,
  LA_PHASE_COUNT = 7
};
void lapack_amd_phase_enable(int on);
void lapack_amd_phase_reset(void);
void lapack_amd_phase_read(double* ms, long* calls, double* flops);

/* Counter-based generator of seeded test matrices (the role DLARNV plays in
 * the reference tests, SRC/dlarnv.f:94-172, with an algorithm of its own):
 * x[k] for k in [0, n) is a pure function of (seed, offset + k), so any
 * block of a matrix can be produced independently and bit-identically on
 * any host.  idist: 1 = uniform(0,1), 2 = uniform(-1,1), 3 = normal(0,1). */
void lapack_amd_dlarnv(int idist, unsigned long long seed, unsigned long long offset, long long n, double* x);

/* Library version string. */
const char* lapack_amd_version(void);

/* ============================ failure reporting =========================== */

/* The multi-workgroup panel kernels of DGETRF/SGETRF and the triangular
 * sweeps of DGETRS/SGETRS wait for each other's workgroups with bounded
 * waits.  When a wait gives up (its workgroups were not all resident, e.g.
 * another process held the CUs), every kernel still runs to its end and:
 *   - getrf leaves A exactly as passed and writes INFO = LA_ERR_SYNC;
 *   - getrs leaves B exactly as passed;
 *   - bit 0 of the device's sticky status word is set.
 * The synchronous entry points (Fortran drop-ins and lapack_amd_*) then redo
 * the call once with forms that never wait (a column-at-a-time panel, a
 * blocked triangular solve) and return its result; the mixed-precision
 * drivers (DSGESV, DGESV_IR, DGERFS) return LA_ERR_SYNC instead.  Callers of
 * the _async entry points read the sticky word here (the call waits for the
 * current device to be idle); clear != 0 resets it.  Returns the word (0 =
 * no wait gave up), or a negative LA_ERR_* code. */
int lapack_amd_device_status(int clear);

/* Test hook: the next `count` factorizations or solves start with their
 * wait-error word already set, so each of their bounded waits gives up at
 * once (the failure path above, forced from the host; no kernel hangs). */
void lapack_amd_test_fail_waits(int count);

#ifdef __cplusplus
}
#endif
#endif /* LAPACK_AMD_H */
