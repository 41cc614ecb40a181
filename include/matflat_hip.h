/*
 * matflat_hip.h -- C ABI of libmatflat_hip.so, the MI355X (gfx950) batched
 * implementation of MatFlat's dense hot path.
 *
 * Conventions shared by every entry point:
 *  - Matrices are column-major; element (i,j) of a matrix with leading
 *    dimension ld sits at a[i + j*ld].  Rows [rows, ld) are never read or
 *    written.
 *  - "vbatched" entry points take device arrays of per-matrix sizes, leading
 *    dimensions and pointers (count entries each).  Every pointer argument
 *    (arrays and the matrices they point to) is DEVICE memory.  max_n is a
 *    host-side upper bound on the sizes; a matrix that exceeds it, or that
 *    has an invalid size or leading dimension, gets a negative info entry
 *    (-(argument position), LAPACK convention) and is left untouched.
 *  - Enum values are the CBLAS integers MatFlat uses (Transpose.cs:8-29,
 *    Uplo.cs:8-19): NoTrans 111, Trans 112, ConjTrans 113, ConjNoTrans 114,
 *    Upper 121, Lower 122.
 *  - Every call is asynchronous on the handle's stream and returns a status:
 *    MFH_OK, or a negative value naming the invalid host argument
 *    (-position), or MFH_ERR_HIP / MFH_ERR_ALLOC / MFH_ERR_UNSUPPORTED.
 *    mfh_last_error() describes the most recent failure on this thread.
 *  - Numerical failures are per matrix in the device info array: 0 = ok,
 *    j+1 = the factorization failed at row j (MatFlat throws
 *    MatrixFactorizationException there, MatrixFactorizationException.cs:9-30).
 *  - A handle is used by one host thread at a time.
 *
 * The FP64 LU, Cholesky, triangular-solve and GEMM kernels reproduce
 * MatFlat's summation order exactly (Internals.Dot, Internals.cs:265-292),
 * so their results are bit-identical to MatFlat's CPU results.
 */
#ifndef MATFLAT_HIP_H
#define MATFLAT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFH_OK 0
#define MFH_ERR_HIP (-1001)
#define MFH_ERR_ALLOC (-1002)
#define MFH_ERR_UNSUPPORTED (-1003)
#define MFH_ERR_STATE (-1004)

#define MFH_NOTRANS 111
#define MFH_TRANS 112
#define MFH_CONJTRANS 113
#define MFH_CONJNOTRANS 114
#define MFH_UPPER 121
#define MFH_LOWER 122
#define MFH_NONUNIT 131
#define MFH_UNIT 132

/* Largest order of the MatFlat-order (exact) LU, Cholesky, SPD solve,
 * triangular solve and LU solve, double and complex. */
#define MFH_EXACT_MAX_N 1024
/* Largest order of the LU inverse and of the SPD-solve adjoint. */
#define MFH_SMALL_MAX_N 128
/* Largest order of the blocked MFMA factorizations (mixed precision). */
#define MFH_BLOCKED_MAX_N 512

typedef struct mfh_handle mfh_handle;
typedef struct mfh_plan mfh_plan;

/* ---- handle ------------------------------------------------------------ */
int mfh_create(mfh_handle** handle);
int mfh_destroy(mfh_handle* handle);
/* stream is a hipStream_t (0 = the null stream). */
int mfh_set_stream(mfh_handle* handle, void* stream);
int mfh_get_stream(mfh_handle* handle, void** stream);
int mfh_synchronize(mfh_handle* handle);
const char* mfh_last_error(void);
/* Library version as major*10000 + minor*100 + patch. */
int mfh_version(void);

/* Math mode of a handle.  MFH_MATH_EXACT (the default) reproduces MatFlat's
 * floating-point operation order, so results are bit-identical to MatFlat's
 * CPU results.  MFH_MATH_FAST lets the SPD factorization / solve kernels use
 * fused multiply-adds and a right-looking, register-blocked schedule; results
 * then agree with MatFlat to rounding (normwise backward error of order
 * n*eps) instead of bit for bit.  No reference counterpart: MatFlat has one
 * (sequential) CPU path. */
#define MFH_MATH_EXACT 0
#define MFH_MATH_FAST 1
int mfh_set_math_mode(mfh_handle* handle, int mode);
int mfh_get_math_mode(mfh_handle* handle, int* mode);

/* Tuning of the FAST Cholesky / SPD-solve kernels (no MatFlat counterpart):
 * matrices of up to 8 * row_blocks rows (n, or n + 1 for the SPD solve;
 * row_blocks 1..17) run on `waves` wavefronts per matrix: 1, 2 or 4 (2 needs
 * row_blocks >= 2, 4 needs row_blocks >= 4), or 0 for the library's built-in
 * table.  matflat_amd/autotune.py measures every choice per bucket. */
int mfh_set_fast_waves(mfh_handle* handle, int row_blocks, int waves);
int mfh_get_fast_waves(mfh_handle* handle, int row_blocks, int* waves);

/* Workgroup shape of the exact (MatFlat-order) Cholesky / SPD-solve kernel
 * for size bucket `bucket` (matrices of 16*bucket+1 .. 16*(bucket+1) rows,
 * n + 1 rows for the SPD solve): `lanes` lanes per matrix (8, 16, 32 or 64;
 * 64 / lanes matrices per wavefront) and `rows_per_lane` rows each (2 or 3),
 * among the built shapes (8,2) (16,2) (16,3) (32,2) (64,2) (64,3) with
 * lanes * rows_per_lane at least the bucket's rows and the packed triangles
 * of a wavefront's matrices within 160 KiB of LDS; (0, 0) restores the
 * built-in choice.  Results do not depend on the shape (every element is
 * the same MatFlat-order dot product); matflat_amd/autotune.py picks the
 * fastest per bucket. */
int mfh_set_chol_shape(mfh_handle* handle, int bucket, int lanes, int rows_per_lane);

/* Concurrency of the FAST-mode Cholesky / SPD-solve size buckets (no MatFlat
 * counterpart): the per-bucket launches are forked over `streams` (1..4,
 * default 2) streams of default priority that the handle owns, largest
 * bucket first, and joined back into the handle's stream, so a bucket's
 * last partial round overlaps the next bucket.  1 keeps them in order on
 * the handle's stream.  With per-launch timing on (mfh_set_timing) the
 * launches always run in order, so each is timed alone. */
int mfh_set_bucket_streams(mfh_handle* handle, int streams);
int mfh_get_bucket_streams(mfh_handle* handle, int* streams);

/* Per-launch kernel timing (a profiling aid; no MatFlat counterpart).  With
 * timing enabled, every kernel launch of the factorization / solve entry
 * points on this handle (outside plan capture) is bracketed by HIP events on
 * the handle's stream.  mfh_get_timing waits for the recorded launches,
 * returns the first min(capacity, *n_records) of them in launch order --
 * kind (MFH_KERNEL_*), rows (the size bucket's row bound: the largest
 * number of rows a matrix of that launch holds, 0 for the size sort) and
 * elapsed milliseconds -- and clears the record.  *n_records receives the
 * number recorded. */
#define MFH_KERNEL_SORT 1
#define MFH_KERNEL_CHOL_EXACT 2
#define MFH_KERNEL_CHOL_FAST 3
#define MFH_KERNEL_GETRF_EXACT 4
#define MFH_KERNEL_TRSV_EXACT 5
#define MFH_KERNEL_GETRS_EXACT 6
#define MFH_KERNEL_GESV_MIXED 7
#define MFH_KERNEL_GESV_F64 8
#define MFH_KERNEL_GESV_STEP 9
#define MFH_KERNEL_GEMV_EXACT 10
#define MFH_KERNEL_GETRI_SETUP 11
#define MFH_KERNEL_GETRI_EXACT 12
#define MFH_KERNEL_GETRF_GEN 13
#define MFH_KERNEL_POTRF_GEN 14
#define MFH_KERNEL_TRSV_GEN 15
#define MFH_KERNEL_GETRS_PERMUTE 16
#define MFH_KERNEL_GEVD_EXACT 17
int mfh_set_timing(mfh_handle* handle, int enable);
int mfh_get_timing(mfh_handle* handle, int capacity, int* n_records, int* kind, int* rows,
                   float* ms);

/* ---- LU with partial pivoting -------------------------------------------
 * Replaces Factorization.Lu(int m, int n, double* a, int lda, int* piv)
 * (Factorization.Lu.cs:122-161, LuCore :296-350) for a batch.  Computes
 * P*A = L*U in place; piv[i] is the original row now at row i (0-based
 * permutation of length m); sign[b] = +-1 is the returned pivot sign.
 * Singular matrices are not an error (MatFlat never throws here).
 * m, n <= MFH_EXACT_MAX_N. */
int mfh_dgetrf_vbatched(mfh_handle* h, int count, const int* m, const int* n,
                        double* const* a, const int* lda, int* const* piv,
                        int* sign, int* info, int max_m, int max_n);
/* Same for count matrices of one shape spaced stride_a doubles apart, with
 * pivots spaced stride_piv ints apart. */
int mfh_dgetrf_strided_batched(mfh_handle* h, int count, int m, int n, double* a, int lda,
                               int64_t stride_a, int* piv, int64_t stride_piv,
                               int* sign, int* info);

/* ---- Cholesky ------------------------------------------------------------
 * Replaces Factorization.Cholesky(int n, double* a, int lda)
 * (Factorization.Cholesky.cs:114-168).  Reads the upper triangle only,
 * writes L to the lower triangle and zeroes the strict upper triangle.  On
 * failure at row j info = j+1 and the matrix holds exactly the partial
 * writes MatFlat leaves when it throws.  n <= MFH_EXACT_MAX_N. */
int mfh_dpotrf_vbatched(mfh_handle* h, int count, const int* n, double* const* a,
                        const int* lda, int* info, int max_n);

/* The SPD solve above, allocating nothing.  No counterpart in the reference:
 * it is for callers that cannot call hipMalloc (code under hipGraph capture,
 * tight serving loops, harnesses that forbid allocation).  Same arguments,
 * results and errors as mfh_dposv_vbatched, for count <= MFH_NOALLOC_MAX_COUNT
 * (a larger count returns -2).  Its scratch (the size sort's 2 ints per
 * matrix and histograms, 8 MiB + 4 KiB) is a static device array of the
 * library, one copy per GPU, reset by a kernel at the start of every call:
 * calls that overlap on one GPU (on different streams or handles) race on
 * it.  It creates no stream or event: the size buckets run in order unless
 * the handle already owns its side streams (from an earlier FAST call).
 * max_n <= MFH_SMALL_MAX_N (larger orders need a workspace; -8 otherwise). */
#define MFH_NOALLOC_MAX_COUNT 1048576
int mfh_dposv_vbatched_noalloc(mfh_handle* h, int count, const int* n, double* const* a,
                               const int* lda, double* const* x, int* info, int max_n);

/* ---- triangular solve ----------------------------------------------------
 * Replaces Blas.SolveTriangular(Uplo, Transpose, int n, double* a, int lda,
 * double* x, int incx) (Blas.SolveTriangular.cs:162-242).  diag MFH_UNIT is
 * the unit-lower/upper variant LuInverse emulates by copying the factor
 * (Factorization.LuInverse.cs:174-182); MFH_NONUNIT is MatFlat's own.  One
 * (uplo, trans, diag) for the whole batch; n <= MFH_EXACT_MAX_N. */
int mfh_dtrsv_vbatched(mfh_handle* h, int uplo, int trans, int diag, int count,
                       const int* n, const double* const* a, const int* lda,
                       double* const* x, const int* incx, int* info, int max_n);

/* ---- SPD solve (Cholesky + two triangular solves, fused) -----------------
 * Per matrix: Factorization.Cholesky, then SolveTriangular(Lower, NoTrans)
 * and SolveTriangular(Lower, Trans) on x (Factorization.Cholesky.cs:114-168,
 * Blas.SolveTriangular.cs:214-231).  x holds b on entry (unit stride).  On
 * a failed factorization x is left unchanged.  n <= MFH_EXACT_MAX_N. */
int mfh_dposv_vbatched(mfh_handle* h, int count, const int* n, double* const* a,
                       const int* lda, double* const* x, int* info, int max_n);

/* ---- LU solve --------------------------------------------------------------
 * The solve composition of Factorization.LuInverse.cs:169-192 for one
 * right-hand side per matrix: y[j] = b[piv[j]]; y <- L1^-1 y (unit lower);
 * x <- U^-1 y (Upper, NoTrans).  lu and piv as produced by
 * mfh_dgetrf_vbatched; b and x may alias.  n <= MFH_EXACT_MAX_N. */
int mfh_dgetrs_vbatched(mfh_handle* h, int count, const int* n, const double* const* lu,
                        const int* lda, const int* const* piv, const double* const* b,
                        double* const* x, int* info, int max_n);

/* ---- LU inverse ------------------------------------------------------------
 * Factorization.LuInverse(int n, double* a, int lda, int* piv)
 * (Factorization.LuInverse.cs:105-192) for a batch: a and piv hold
 * mfh_dgetrf_vbatched's output, and a is overwritten by A^-1, bit-identical
 * to MatFlat (column j is the LU-solve composition above applied to e_j).
 * A matrix whose U has a zero diagonal entry gets info = j + 1 for the first
 * such j and is left unchanged (MatFlat throws "The matrix is not
 * invertible." before writing).  Padding rows are never touched.
 * n <= MFH_SMALL_MAX_N; count * max_n < 2^31.  Staging lives in a second
 * handle buffer of about count * max_n^2 * 8 bytes. */
int mfh_dgetri_vbatched(mfh_handle* h, int count, const int* n, double* const* a, const int* lda,
                        const int* const* piv, int* info, int max_n);

/* ---- GEMM ------------------------------------------------------------------
 * Replaces Blas.MulMatMat(Transpose, Transpose, m, n, k, a, lda, b, ldb, c,
 * ldc) (Blas.MulMatMat.cs:226-368 double, :407-648 Complex) for a batch with
 * per-product transposes, extended to C = alpha*op(A)*op(B) + beta*C.  C is
 * not read when beta == 0; with alpha = 1, beta = 0 the result is exactly
 * MatFlat's.  A product whose c pointer equals its a or b pointer gets
 * info = -1 (MatFlat's "In-place operation is not supported").
 * m, n, k <= 4096 (larger k is fine; sizes are not staged whole).
 * alpha and beta are device arrays of count doubles, one per product. */
int mfh_dgemm_vbatched(mfh_handle* h, int count, const int* transa, const int* transb,
                       const int* m, const int* n, const int* k, const double* alpha,
                       const double* const* a, const int* lda, const double* const* b,
                       const int* ldb, const double* beta, double* const* c, const int* ldc,
                       int* info, int max_m, int max_n, int max_k);
/* Blas.MulMatVec(Transpose trans, int m, int n, double* a, int lda, double* x,
 * int incx, double* y, int incy) (Blas.MulMatVec.cs:138-206) for a batch
 * with per-product transposes: y = A x (NoTrans, m outputs) or y = A^T x
 * (Trans, n outputs), every output one Internals.Dot in MatFlat's order, so
 * y is bit-identical to the reference.  Per-product checks follow the
 * reference's order and set info = -(position in this argument list); a
 * product whose y pointer equals its a or x pointer gets info = -1
 * ("In-place operation is not supported"); the transpose is checked last,
 * as in the reference.  lda * n, incx * len(x) and incy * len(y) must stay
 * below 2^28 elements; m <= max_m, n <= max_n. */
int mfh_dgemv_vbatched(mfh_handle* h, int count, const int* trans, const int* m, const int* n,
                       const double* const* a, const int* lda, const double* const* x,
                       const int* incx, double* const* y, const int* incy, int* info,
                       int max_m, int max_n);
/* Complex version: a, b and c hold interleaved (re, im) doubles (the layout
 * of System.Numerics.Complex and hipDoubleComplex); alpha and beta are device
 * arrays of count (re, im) pairs, one per product; transa/transb may also be
 * MFH_CONJTRANS (113) or MFH_CONJNOTRANS (114). */
int mfh_zgemm_vbatched(mfh_handle* h, int count, const int* transa, const int* transb,
                       const int* m, const int* n, const int* k, const double* alpha,
                       const double* const* a, const int* lda, const double* const* b,
                       const int* ldb, const double* beta, double* const* c, const int* ldc,
                       int* info, int max_m, int max_n, int max_k);

/* ---- complex128 LU, Cholesky and triangular solve ----------------------------
 * a and x hold interleaved (re, im) doubles (System.Numerics.Complex's
 * layout).  Bit-identical to MatFlat: products, sums and divisions restate
 * the BCL's Complex operators (division by Smith's method), dot products
 * keep Internals.Dot / DotConj's pairing.
 *
 * Factorization.Lu(int m, int n, Complex* a, int lda, int* piv)
 * (Factorization.Lu.cs:199-238, LuCore :352-406): as mfh_dgetrf_vbatched,
 * with the pivot chosen by |re| + |im| (Internals.FastMagnitude,
 * Internals.cs:15-19).  m, n <= MFH_EXACT_MAX_N. */
int mfh_zgetrf_vbatched(mfh_handle* h, int count, const int* m, const int* n,
                        double* const* a, const int* lda, int* const* piv,
                        int* sign, int* info, int max_m, int max_n);
/* Factorization.Cholesky(int n, Complex* a, int lda)
 * (Factorization.Cholesky.cs:195-249): A = L L^H from the upper triangle
 * (only the real part of the diagonal is read); L's diagonal is real.  On
 * failure at row j, info = j+1 and the matrix holds MatFlat's partial
 * writes.  n <= MFH_EXACT_MAX_N. */
int mfh_zpotrf_vbatched(mfh_handle* h, int count, const int* n, double* const* a,
                        const int* lda, int* info, int max_n);
/* Blas.SolveTriangular(Uplo, Transpose, int n, Complex* a, int lda,
 * Complex* x, int incx) (Blas.SolveTriangular.cs:280-395, MulSub :397-421,
 * MulSubConj :423-445) for a batch; trans is any of MFH_NOTRANS, MFH_TRANS,
 * MFH_CONJTRANS, MFH_CONJNOTRANS (the Conj forms use conj(A)).  Non-unit
 * diagonal, as the reference.  One (uplo, trans) for the whole batch;
 * n <= MFH_EXACT_MAX_N. */
int mfh_ztrsv_vbatched(mfh_handle* h, int uplo, int trans, int count, const int* n,
                       const double* const* a, const int* lda, double* const* x,
                       const int* incx, int* info, int max_n);

/* ---- mixed-precision LU solve with FP64 iterative refinement --------------
 * No single MatFlat counterpart: it composes Factorization.Lu and the LU
 * solve of LuInverse (above) in lower precision and refines in FP64.
 * Factors each A (FP64, n <= MFH_BLOCKED_MAX_N) in FP32 with partial
 * pivoting (BF16x6 split MFMA trailing update when n >= 256, FP32
 * accumulation), then iterates x <- x + A_lo^-1 (b - A x) with the residual,
 * the update and the test in FP64 until
 * ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= tol_per_n * n or
 * max_iter refinements.  iters[b] = refinements taken, berr[b] = final
 * backward error; info[b] = 0, j+1 for an exactly zero pivot in column j,
 * n+1 if the backward error did not reach the tolerance (a NaN backward
 * error counts as not reached), or n+2 (iters[b] = -2, x untouched) if an
 * entry of A is NaN or beyond the FP32 range -- such a matrix cannot be
 * factored in FP32 (LAPACK dsgesv's check) and is for mfh_dgesv_vbatched.
 * A and b are not modified; x receives the solution. */
int mfh_dgesv_mixed_vbatched(mfh_handle* h, int count, const int* n, const double* const* a,
                             const int* lda, const double* const* b, double* const* x,
                             int max_iter, double tol_per_n, int* iters, double* berr,
                             int* info, int max_n);
/* Plain FP64 blocked LU solve for the same inputs (the comparison path of
 * the mixed-precision solver), with berr as above; info[b] = n+1 when the
 * backward error is not finite. */
int mfh_dgesv_vbatched(mfh_handle* h, int count, const int* n, const double* const* a,
                       const int* lda, const double* const* b, double* const* x,
                       double* berr, int* info, int max_n);
/* The factorization of that solver on its own: Factorization.Lu
 * (Factorization.Lu.cs:122-161) for n <= MFH_BLOCKED_MAX_N in FAST math --
 * blocked right-looking partial pivoting with the FP64 trailing update on
 * matrix cores.  P*A = L*U in place (unit lower L below the diagonal, U on
 * and above it) and piv[i] = the original row now at row i, as MatFlat
 * returns them.  The pivot of a column is its largest |a| to 2^-20 relative
 * (|a| compared in FP32), lowest row on ties, so on near-ties piv may differ
 * from MatFlat's order; the factors meet ||PA - LU|| / (||A|| n eps) <= 10.
 * info[b] = j+1 for the first exactly zero pivot (MatFlat does not fail
 * there either: the factorization completes), else 0.  Square matrices only;
 * piv[b] holds n ints. */
int mfh_dgetrf_blocked_vbatched(mfh_handle* h, int count, const int* n, double* const* a,
                                const int* lda, int* const* piv, int* info, int max_n);

/* ---- adjoint of the SPD solve ----------------------------------------------
 * No MatFlat counterpart (MatFlat has no derivatives).  For x = A^-1 b with
 * A = L L^T (l from mfh_dposv_vbatched / mfh_dpotrf_vbatched) and
 * xbar = dloss/dx: bbar = A^-1 xbar and
 * abar = -(bbar x^T + x bbar^T)/2, written as a full symmetric n x n matrix
 * with leading dimension ldabar.  n <= MFH_SMALL_MAX_N. */
int mfh_dposv_backward_vbatched(mfh_handle* h, int count, const int* n, const double* const* l,
                                const int* lda, const double* const* x, const double* const* xbar,
                                double* const* abar, const int* ldabar, double* const* bbar,
                                int* info, int max_n);

/* ---- adjoint of the Cholesky factorization ---------------------------------
 * No MatFlat counterpart.  For A = L L^T (l from mfh_dpotrf_vbatched, lower
 * triangle read) and lbar = dloss/dL (lower triangle read, the strict upper
 * ignored): abar = L^-T M L^-1, M = (Phi(L^T Lbar) + Phi(L^T Lbar)^T) / 2 with
 * Phi = the lower triangle with its diagonal halved -- the gradient with
 * respect to the symmetric matrix the factor came from, written as a full
 * symmetric n x n matrix with leading dimension ldabar (which must not
 * overlap l or lbar).  Both triangular solves run in dot form with
 * Internals.Dot's pairing (Internals.cs:265-292).  info[b] = 0, or -(position) for invalid
 * per-matrix arguments.  No check of L's diagonal: a zero gives inf/NaN, as
 * the triangular solve does.  n <= MFH_SMALL_MAX_N. */
int mfh_dpotrf_backward_vbatched(mfh_handle* h, int count, const int* n, const double* const* l,
                                 const int* ldl, const double* const* lbar, const int* ldlbar,
                                 double* const* abar, const int* ldabar, int* info, int max_n);

/* ---- generalized symmetric-definite eigenproblem ---------------------------
 * Factorization.Gevd (double), Factorization.Gevd.cs:145-186 with GevdCore
 * :330-376: A v = w B v for symmetric A and symmetric positive definite B,
 * both read from their upper triangles only.  In MatFlat's operation order
 * (bit-identical): b is overwritten by the Cholesky factor L (B = L L^T,
 * strict upper triangle zeroed), w by the n eigenvalues in descending order
 * and a by the eigenvectors (columns, B-orthonormal).  info[b] = 0; j+1 when
 * B is not positive definite at row j (MatFlat throws MatFlatException; b
 * holds the Cholesky's partial writes, a and w are untouched); n+1 when the
 * singular-value iteration does not converge (MatFlat throws
 * MatrixFactorizationException; a holds the vectors reached, w untouched);
 * -(position) for invalid per-matrix arguments (nothing written).  Orders up
 * to MFH_SMALL_MAX_N; the math mode does not apply (always MatFlat's order). */
int mfh_dgevd_vbatched(mfh_handle* h, int count, const int* n, double* const* a, const int* lda,
                       double* const* b, const int* ldb, double* const* w, int* info, int max_n);

/* ---- plans: one batch captured into a hipGraph ------------------------------
 * A plan records one call of an entry point with the given arguments into a
 * hipGraph (device arrays stay owned by the caller and must outlive the
 * plan) and replays it with mfh_plan_execute on the handle's stream, with
 * no host work: the size sort, every size-bucket launch (forked over the
 * handle's side streams as graph edges) and the solve steps are nodes of the
 * graph.  The plan owns the workspace the call needs, so later calls on the
 * handle do not disturb it.  Replays read the per-matrix arrays afresh:
 * sizes and data may change between replays within the captured maxima.
 * The create functions take the arguments of the entry point they name
 * (no counterpart in the reference, which has no batches) followed by the
 * plan; mfh_plan_get_op returns the MFH_OP_* of a plan. */
#define MFH_OP_DPOSV 1
#define MFH_OP_DGETRF 2
#define MFH_OP_DPOTRF 3
#define MFH_OP_DTRSV 4
#define MFH_OP_DGESV 5
#define MFH_OP_DGESV_MIXED 6
#define MFH_OP_DGEMM 7
#define MFH_OP_ZGEMM 8
int mfh_plan_create_dposv(mfh_handle* h, int count, const int* n, double* const* a,
                          const int* lda, double* const* x, int* info, int max_n,
                          mfh_plan** plan);
int mfh_plan_create_dgetrf(mfh_handle* h, int count, const int* m, const int* n,
                           double* const* a, const int* lda, int* const* piv, int* sign,
                           int* info, int max_m, int max_n, mfh_plan** plan);
int mfh_plan_create_dpotrf(mfh_handle* h, int count, const int* n, double* const* a,
                           const int* lda, int* info, int max_n, mfh_plan** plan);
int mfh_plan_create_dtrsv(mfh_handle* h, int uplo, int trans, int diag, int count,
                          const int* n, const double* const* a, const int* lda,
                          double* const* x, const int* incx, int* info, int max_n,
                          mfh_plan** plan);
int mfh_plan_create_dgesv(mfh_handle* h, int count, const int* n, const double* const* a,
                          const int* lda, const double* const* b, double* const* x,
                          double* berr, int* info, int max_n, mfh_plan** plan);
int mfh_plan_create_dgesv_mixed(mfh_handle* h, int count, const int* n,
                                const double* const* a, const int* lda,
                                const double* const* b, double* const* x, int max_iter,
                                double tol_per_n, int* iters, double* berr, int* info,
                                int max_n, mfh_plan** plan);
/* The batched GEMM of mfh_dgemm_vbatched / mfh_zgemm_vbatched (same
 * arguments) captured into one graph: tile-list build and the GEMM kernel,
 * in the handle's math mode at capture time.  Replays read the per-product
 * arrays afresh, so sizes and data may change between replays within the
 * captured max_m and max_n. */
int mfh_plan_create_dgemm(mfh_handle* h, int count, const int* transa, const int* transb,
                          const int* m, const int* n, const int* k, const double* alpha,
                          const double* const* a, const int* lda, const double* const* b,
                          const int* ldb, const double* beta, double* const* c, const int* ldc,
                          int* info, int max_m, int max_n, int max_k, mfh_plan** plan);
int mfh_plan_create_zgemm(mfh_handle* h, int count, const int* transa, const int* transb,
                          const int* m, const int* n, const int* k, const double* alpha,
                          const double* const* a, const int* lda, const double* const* b,
                          const int* ldb, const double* beta, double* const* c, const int* ldc,
                          int* info, int max_m, int max_n, int max_k, mfh_plan** plan);
int mfh_plan_get_op(mfh_plan* plan, int* op);
int mfh_plan_execute(mfh_plan* plan);
int mfh_plan_destroy(mfh_plan* plan);

#ifdef __cplusplus
}
#endif
#endif
