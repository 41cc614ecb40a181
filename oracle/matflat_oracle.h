/*
 * matflat_oracle.h -- CPU restatement of MatFlat's hot-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * kernels in matflat_amd/csrc.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it; the product path never
 * calls it and has no CPU fallback.
 *
 * Every function follows the operation order of the cited reference
 * (MatFlat 0.8.2, C#) so that IEEE binary64 results are bit-identical when
 * compiled with -O2 -ffp-contract=off -fno-fast-math.  Matrices are
 * column-major, element (i,j) at a[i + j*lda].  Complex numbers are
 * interleaved (re, im) doubles.
 */
#ifndef MATFLAT_ORACLE_H
#define MATFLAT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CBLAS enum values (Transpose.cs:8-29, Uplo.cs:8-19). */
enum { MFO_NOTRANS = 111, MFO_TRANS = 112, MFO_CONJTRANS = 113, MFO_CONJNOTRANS = 114 };
enum { MFO_UPPER = 121, MFO_LOWER = 122 };

/* Internals.Dot(int n, double* x, int incx, double* y, int incy), Internals.cs:265-292. */
double mfo_ddot(int n, const double* x, int incx, const double* y, int incy);

/* Factorization.Lu (LuCore double), Factorization.Lu.cs:296-350.  Returns the pivot sign. */
int mfo_dgetrf(int m, int n, double* a, int lda, int* piv);

/* Factorization.Cholesky double, Factorization.Cholesky.cs:114-168.
 * Returns 0, or j+1 for the first row whose diagonal is not > 0 (the
 * reference throws there; the matrix then holds the same partial writes). */
int mfo_dpotrf(int n, double* a, int lda);

/* Blas.SolveTriangular double, Blas.SolveTriangular.cs:162-242. */
int mfo_dtrsv(int uplo, int trans, int n, const double* a, int lda, double* x, int incx);

/* Unit-diagonal lower solve that LuInverse emulates by copying L with a unit
 * diagonal (Factorization.LuInverse.cs:174-182): x_i = (x_i - dot)/1.0. */
void mfo_dtrsv_unit_lower(int n, const double* a, int lda, double* x, int incx);

/* LU solve composition of Factorization.LuInverse.cs:169-192 for one
 * right-hand side: y[j] = b[piv[j]]; y <- L1^-1 y; x <- U^-1 y. */
void mfo_dgetrs(int n, const double* lu, int lda, const int* piv, const double* b, double* x);

/* Factorization.LuInverse double, Factorization.LuInverse.cs:105-192: a and
 * piv hold Factorization.Lu's output; a is overwritten by A^-1.  Returns
 * j+1 for the first zero diagonal entry of U (the reference throws
 * MatrixFactorizationException before writing anything), else 0. */
int mfo_dgetri(int n, double* a, int lda, const int* piv);

/* Cholesky followed by Lower/NoTrans and Lower/Trans solves. Returns info. */
int mfo_dposv(int n, double* a, int lda, double* x);

/* Blas.MulMatMat double, Blas.MulMatMat.cs:226-368, extended by
 * C = alpha*dot + beta*C0 (C0 not read when beta == 0).  With alpha=1,
 * beta=0 this is exactly the reference. */
int mfo_dgemm(int transa, int transb, int m, int n, int k, double alpha,
              const double* a, int lda, const double* b, int ldb,
              double beta, double* c, int ldc);

/* Blas.MulMatVec double, Blas.MulMatVec.cs:138-206: NoTrans gives m outputs
 * y[i*incy] = Dot(n, a + i, lda, x, incx) (row i of A); Trans gives n
 * outputs y[i*incy] = Dot(m, a + i*lda, 1, x, incx) (column i).  Returns -1
 * for a transpose other than NoTrans/Trans (the reference's last check). */
int mfo_dgemv(int trans, int m, int n, const double* a, int lda, const double* x, int incx,
              double* y, int incy);

/* Blas.MulMatMat Complex, Blas.MulMatMat.cs:407-575 with Dot_NN/CN/NC/CC
 * (:598-648), same alpha/beta extension (complex alpha, beta). */
int mfo_zgemm(int transa, int transb, int m, int n, int k, const double* alpha,
              const double* a, int lda, const double* b, int ldb,
              const double* beta, double* c, int ldc);

/* ---- complex128 (interleaved re, im doubles; element (i, j) at 2(i + j*lda)) */

/* BCL Complex division (Smith's method, branch on |d| < |c|):
 * (a + bi) / (c + di). */
void mfo_zdiv(double a, double b, double c, double d, double* re, double* im);

/* Internals.Dot / Internals.DotConj over Complex, Internals.cs:294-321 and
 * :352-380 (conj = 1: products conj(x)*y); out = (re, im). */
void mfo_zdot(int conj, int n, const double* x, int incx, const double* y, int incy, double* out);

/* Factorization.Lu Complex, Factorization.Lu.cs:199-238 / :352-406: pivot by
 * |re| + |im|; returns the pivot sign, piv as mfo_dgetrf. */
int mfo_zgetrf(int m, int n, double* a, int lda, int* piv);

/* Factorization.Cholesky Complex, Factorization.Cholesky.cs:195-249: A = L L^H
 * from the upper triangle; returns 0 or j + 1 of the failing row (partial
 * writes as the reference leaves them). */
int mfo_zpotrf(int n, double* a, int lda);

/* Blas.SolveTriangular Complex, Blas.SolveTriangular.cs:280-395: every
 * uplo / transpose (NoTrans, Trans, ConjTrans, ConjNoTrans). */
int mfo_ztrsv(int uplo, int trans, int n, const double* a, int lda, double* x, int incx);

/* ---- generalized symmetric-definite eigenproblem -------------------------- */

/* Internals.Norm(int n, double* x), Internals.cs:435-460. */
double mfo_dnorm(int n, const double* x);

/* Internals.Drotg, Internals.cs:626-668 (reference BLAS drotg). */
void mfo_drotg(double* da, double* db, double* c, double* s);

/* SvdCore double (Factorization.Svd.cs:786-1256) for a square matrix with
 * left vectors only: s descending, u (ldu) the left singular vectors; a is
 * destroyed; work, e, stmp hold n doubles each.  Returns 0, or 1 when the
 * iteration did not converge (the reference throws). */
int mfo_dsvd_u(int n, double* a, int lda, double* s, double* u, int ldu, double* work, double* e,
               double* stmp);

/* Factorization.Gevd double, Factorization.Gevd.cs:145-186 / :330-376: A v = w B v
 * for symmetric A (upper triangle read) and SPD B (upper triangle read).  b
 * gets L (B = L L^T), w the eigenvalues (descending), a the eigenvectors
 * (L^-T U).  c holds n*n doubles, work 3n.  Returns 0, j+1 if B's Cholesky
 * fails at row j, n+1 if the SVD did not converge. */
int mfo_dgevd(int n, double* a, int lda, double* b, int ldb, double* w, double* c, double* work);

/* ---- batched drivers (CPU baseline; OpenMP over matrices) ---------------- */

/* Cholesky + two solves over a batch stored back-to-back: matrix i at
 * a + off[i], order n[i], leading dimension lda[i]; rhs at x + xoff[i].
 * Parallel over matrices with OpenMP dynamic scheduling when threads > 1. */
void mfo_dposv_batch(int count, const int* n, const int* lda, const int64_t* off,
                     double* a, const int64_t* xoff, double* x, int* info, int threads);

/* LU over a fixed-size batch (stride elements apart). */
void mfo_dgetrf_batch(int count, int m, int n, int lda, int64_t stride, double* a,
                      int* piv, int* sign, int threads);

/* mfo_dgetrf + mfo_dgetrs per matrix over a fixed-size batch (stride
 * elements apart), right-hand sides and solutions n apart. */
void mfo_dgesv_batch(int count, int n, int lda, int64_t stride, double* a, int* piv,
                     const double* b, double* x, int threads);

/* mfo_dgemm (cplx = 0) or mfo_zgemm (cplx = 1) over a batch; offsets count
 * doubles, complex alpha/beta are (re, im) pairs.  info[i] = the call's
 * return value. */
void mfo_gemm_vbatch(int cplx, int count, const int* ta, const int* tb, const int* m, const int* n,
                     const int* k, const double* alpha, const double* a, const int64_t* offa,
                     const int* lda, const double* b, const int64_t* offb, const int* ldb,
                     const double* beta, double* c, const int64_t* offc, const int* ldc, int* info,
                     int threads);

/* mfo_dgevd over a batch stored back-to-back: A, B and w of matrix i at
 * a + aoff[i], b + boff[i], w + woff[i]; info[i] = its return value. */
void mfo_dgevd_batch(int count, const int* n, const int* lda, const int* ldb, const int64_t* aoff,
                     const int64_t* boff, const int64_t* woff, double* a, double* b, double* w,
                     int* info, int threads);

/* mfo_dposv_backward over a fixed-size batch (L and Abar stride elements
 * apart with leading dimension lda; x, w and bbar n apart). */
void mfo_dposv_backward_batch(int count, int n, int lda, int64_t stride, const double* l,
                              const double* x, const double* w, double* abar, double* bbar,
                              int threads);

/* Adjoint of x = A^-1 b with A = L L^T: given L (lower, as produced by
 * mfo_dpotrf), x and w = dloss/dx, returns bbar = A^-1 w and
 * Abar = -(bbar x^T + x bbar^T)/2 (full symmetric n x n, leading dim ldab). */
void mfo_dposv_backward(int n, const double* l, int lda, const double* x, const double* w,
                        double* abar, int ldab, double* bbar);

/* Adjoint of A = L L^T: Abar = L^-T M L^-1 with M the symmetrized
 * Phi(L^T Lbar) / 2 (full symmetric n x n, leading dimension ldab); the
 * operation order is fixed in matflat_oracle.c. */
void mfo_dpotrf_backward(int n, const double* l, int ldl, const double* lbar, int ldlb, double* abar, int ldab);

/* mfo_dpotrf_backward over a fixed-size batch (matrices stride elements apart,
 * leading dimension ld). */
void mfo_dpotrf_backward_batch(int count, int n, int ld, int64_t stride, const double* l, const double* lbar,
                               double* abar, int threads);

#ifdef __cplusplus
}
#endif
#endif
