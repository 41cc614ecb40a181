/*
 * matflat_oracle.c -- CPU restatement of MatFlat 0.8.2's hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels, never the
 * product path.  Build with -O2 -ffp-contract=off -fno-fast-math (see
 * oracle/Makefile) so that every + - * / sqrt below is one IEEE binary64
 * operation in the order written, as RyuJIT executes the C# source.
 *
 * Citations are to files of the reference (MatFlat/...).
 */
#include "matflat_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Internals.Dot(int n, double* x, int incx, double* y, int incy),
 * Internals.cs:265-292.  Odd n: sum = x0*y0 first; even n: sum = 0.0.  Then
 * pairs: sum += x[0]*y[0] + x[incx]*y[incy], i.e. sum + (p0 + p1). */
double mfo_ddot(int n, const double* x, int incx, const double* y, int incy)
{
    double sum;
    if (n & 1) {
        sum = x[0] * y[0];
        x += incx;
        y += incy;
        n--;
    } else {
        sum = 0.0;
    }
    while (n > 0) {
        double p0 = x[0] * y[0];
        double p1 = x[incx] * y[incy];
        sum = sum + (p0 + p1);
        x += 2 * incx;
        y += 2 * incy;
        n -= 2;
    }
    return sum;
}

static inline int imin(int a, int b) { return a < b ? a : b; }

/* Factorization.Lu.cs:296-350 (LuCore, double).  The reference reads
 * colj[j] before testing j < m (Lu.cs:340-341), an out-of-bounds read when
 * n > m; the guard is applied first here, with the same result. */
int mfo_dgetrf(int m, int n, double* a, int lda, int* piv)
{
    double* work = (double*)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < m; i++) piv[i] = i;
    int sign = 1;
    for (int j = 0; j < n; j++) {
        double* colj = a + (size_t)j * lda;
        memcpy(work, colj, sizeof(double) * (size_t)m);
        for (int i = 0; i < m; i++) {
            double s = mfo_ddot(imin(i, j), a + i, lda, work, 1);
            work[i] = work[i] - s;
            colj[i] = work[i];
        }
        int p = j;
        for (int i = j + 1; i < m; i++) {
            if (fabs(work[i]) > fabs(work[p])) p = i;
        }
        if (p != j) {
            /* Internals.SwapRows(n, a + p, a + j, lda), Internals.cs:567-576 */
            for (int c = 0; c < n; c++) {
                double t = a[p + (size_t)c * lda];
                a[p + (size_t)c * lda] = a[j + (size_t)c * lda];
                a[j + (size_t)c * lda] = t;
            }
            int t = piv[p];
            piv[p] = piv[j];
            piv[j] = t;
            sign = -sign;
        }
        if (j < m) {
            double diag = colj[j];
            if (diag != 0.0) {
                /* Internals.DivInplace, Internals.cs:125-147: true division */
                for (int i = j + 1; i < m; i++) colj[i] = colj[i] / diag;
            }
        }
    }
    free(work);
    return sign;
}

/* Factorization.Cholesky.cs:114-168 (double). */
int mfo_dpotrf(int n, double* a, int lda)
{
    for (int j = 0; j < n; j++) {
        double* colj = a + (size_t)j * lda;
        double s = 0.0;
        for (int k = 0; k < j; k++) {
            double* colk = a + (size_t)k * lda;
            double t = (colj[k] - mfo_ddot(k, a + j, lda, a + k, lda)) / colk[k];
            colk[j] = t;
            s = s + t * t;
        }
        s = colj[j] - s;
        if (s > 0) {
            colj[j] = sqrt(s);
        } else {
            return j + 1;
        }
    }
    for (int j = 1; j < n; j++) {
        double* colj = a + (size_t)j * lda;
        for (int i = 0; i < j; i++) colj[i] = 0.0;
    }
    return 0;
}

/* Blas.SolveTriangular.cs:162-242 (double) with MulSub (:397-421):
 * dst[i] -= x[i]*y, element-wise. */
int mfo_dtrsv(int uplo, int trans, int n, const double* a, int lda, double* x, int incx)
{
    if (uplo == MFO_UPPER) {
        if (trans == MFO_NOTRANS) {
            for (int i = n - 1; i >= 0; i--) {
                int p = incx * i;
                const double* acol = a + (size_t)lda * i;
                x[p] = x[p] / acol[i];
                double y = x[p];
                for (int r = 0; r < i; r++) x[r * incx] = x[r * incx] - acol[r] * y;
            }
        } else if (trans == MFO_TRANS) {
            for (int j = 0; j < n; j++) {
                int p = incx * j;
                x[p] = (x[p] - mfo_ddot(j, a + (size_t)lda * j, 1, x, incx)) / a[(size_t)lda * j + j];
            }
        } else {
            return -2;
        }
    } else if (uplo == MFO_LOWER) {
        if (trans == MFO_NOTRANS) {
            for (int i = 0; i < n; i++) {
                int p = incx * i;
                x[p] = (x[p] - mfo_ddot(i, a + i, lda, x, incx)) / a[(size_t)lda * i + i];
            }
        } else if (trans == MFO_TRANS) {
            for (int j = n - 1; j >= 0; j--) {
                int p = incx * j;
                x[p] = x[p] / a[(size_t)j * lda + j];
                double y = x[p];
                for (int r = 0; r < j; r++) x[r * incx] = x[r * incx] - a[j + (size_t)r * lda] * y;
            }
        } else {
            return -2;
        }
    } else {
        return -1;
    }
    return 0;
}

void mfo_dtrsv_unit_lower(int n, const double* a, int lda, double* x, int incx)
{
    for (int i = 0; i < n; i++) {
        int p = incx * i;
        x[p] = (x[p] - mfo_ddot(i, a + i, lda, x, incx)) / 1.0;
    }
}

/* Factorization.LuInverse.cs:169-192 applied to one column: the column of
 * the permutation is replaced by the right-hand side. */
void mfo_dgetrs(int n, const double* lu, int lda, const int* piv, const double* b, double* x)
{
    for (int j = 0; j < n; j++) x[j] = b[piv[j]];
    mfo_dtrsv_unit_lower(n, lu, lda, x, 1);
    mfo_dtrsv(MFO_UPPER, MFO_NOTRANS, n, lu, lda, x, 1);
}

int mfo_dposv(int n, double* a, int lda, double* x)
{
    int info = mfo_dpotrf(n, a, lda);
    if (info != 0) return info;
    mfo_dtrsv(MFO_LOWER, MFO_NOTRANS, n, a, lda, x, 1);
    mfo_dtrsv(MFO_LOWER, MFO_TRANS, n, a, lda, x, 1);
    return 0;
}

/* Blas.MulMatMat.cs:226-368 (double).  Every C element is one
 * Internals.Dot of length k; the four transpose cases only change strides. */
int mfo_dgemm(int transa, int transb, int m, int n, int k, double alpha,
              const double* a, int lda, const double* b, int ldb,
              double beta, double* c, int ldc)
{
    if ((transa != MFO_NOTRANS && transa != MFO_TRANS) || (transb != MFO_NOTRANS && transb != MFO_TRANS))
        return -1;
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < m; i++) {
            const double* x;
            int incx;
            const double* y;
            int incy;
            if (transa == MFO_NOTRANS) { x = a + i; incx = lda; }
            else { x = a + (size_t)i * lda; incx = 1; }
            if (transb == MFO_NOTRANS) { y = b + (size_t)j * ldb; incy = 1; }
            else { y = b + j; incy = ldb; }
            double d = mfo_ddot(k, x, incx, y, incy);
            double t = (alpha == 1.0) ? d : alpha * d;
            double* cij = c + i + (size_t)j * ldc;
            if (beta == 0.0) *cij = t;
            else *cij = t + beta * *cij;
        }
    }
    return 0;
}

/* Factorization.LuInverse.cs:127-192: check the diagonal (:127-134); copy
 * the factors to work (ld n) and clear A's columns (:141-149); X = P with
 * X[j, piv[j]] = 1 (:151-154); a Lower/NoTrans solve of every column with
 * the copy's diagonal set to 1 (:156-164); the diagonal restored and an
 * Upper/NoTrans solve of every column (:166-174). */
int mfo_dgetri(int n, double* a, int lda, const int* piv)
{
    for (int j = 0; j < n; j++)
        if (a[(size_t)lda * j + j] == 0.0) return j + 1;
    double* work = (double*)malloc(sizeof(double) * ((size_t)n * n + n));
    double* diag = work + (size_t)n * n;
    for (int j = 0; j < n; j++) {
        double* colj = a + (size_t)lda * j;
        memcpy(work + (size_t)n * j, colj, sizeof(double) * n);
        diag[j] = colj[j];
        memset(colj, 0, sizeof(double) * n);
    }
    for (int j = 0; j < n; j++) a[(size_t)lda * piv[j] + j] = 1.0;
    for (int j = 0; j < n; j++) work[(size_t)n * j + j] = 1.0;
    for (int j = 0; j < n; j++) mfo_dtrsv(MFO_LOWER, MFO_NOTRANS, n, work, n, a + (size_t)lda * j, 1);
    for (int j = 0; j < n; j++) work[(size_t)n * j + j] = diag[j];
    for (int j = 0; j < n; j++) mfo_dtrsv(MFO_UPPER, MFO_NOTRANS, n, work, n, a + (size_t)lda * j, 1);
    free(work);
    return 0;
}

/* Blas.MulMatVec.cs:185-202: one Internals.Dot per output, in output order. */
int mfo_dgemv(int trans, int m, int n, const double* a, int lda, const double* x, int incx,
              double* y, int incy)
{
    if (trans == MFO_NOTRANS) {
        for (int i = 0; i < m; i++) y[(size_t)i * incy] = mfo_ddot(n, a + i, lda, x, incx);
        return 0;
    }
    if (trans == MFO_TRANS) {
        for (int i = 0; i < n; i++) y[(size_t)i * incy] = mfo_ddot(m, a + (size_t)i * lda, 1, x, incx);
        return 0;
    }
    return -1;
}

/* System.Numerics.Complex operator * (BCL): re = a*c - b*d, im = b*c + a*d. */
static inline void cmul(double a, double b, double c, double d, double* re, double* im)
{
    *re = a * c - b * d;
    *im = b * c + a * d;
}

/* Internals.MulConj (Internals.cs:21-29): conj(x)*y written as
 * a = x.re, b = -x.im, c = y.re, d = y.im; re = a*c - b*d, im = a*d + b*c. */
static inline void cmulconj(const double* x, const double* y, double* re, double* im)
{
    double a = x[0], b = -x[1], c = y[0], d = y[1];
    *re = a * c - b * d;
    *im = a * d + b * c;
}

/* Internals.MulConjConj (Internals.cs:31-39): conj(x)*conj(y). */
static inline void cmulconjconj(const double* x, const double* y, double* re, double* im)
{
    double a = x[0], b = -x[1], c = y[0], d = -y[1];
    *re = a * c - b * d;
    *im = a * d + b * c;
}

/* Dot_NN/CN/NC/CC, Blas.MulMatMat.cs:598-648: sequential sum from
 * Complex.Zero, one complex product added per step. */
static void zdot(int kind, int n, const double* x, int incx, const double* y, int incy,
                 double* sre, double* sim)
{
    double sr = 0.0, si = 0.0;
    while (n > 0) {
        double pr, pi;
        switch (kind) {
        case 0: cmul(x[0], x[1], y[0], y[1], &pr, &pi); break;
        case 1: cmulconj(x, y, &pr, &pi); break;     /* Dot_CN */
        case 2: cmulconj(y, x, &pr, &pi); break;     /* Dot_NC */
        default: cmulconjconj(x, y, &pr, &pi); break; /* Dot_CC */
        }
        sr = sr + pr;
        si = si + pi;
        x += 2 * (size_t)incx;
        y += 2 * (size_t)incy;
        n--;
    }
    *sre = sr;
    *sim = si;
}

static int has_conj(int t) { return t == MFO_CONJNOTRANS || t == MFO_CONJTRANS; }
static int remove_conj(int t)
{
    if (t == MFO_CONJNOTRANS) return MFO_NOTRANS;
    if (t == MFO_CONJTRANS) return MFO_TRANS;
    return t;
}

int mfo_zgemm(int transa, int transb, int m, int n, int k, const double* alpha,
              const double* a, int lda, const double* b, int ldb,
              const double* beta, double* c, int ldc)
{
    if (transa < MFO_NOTRANS || transa > MFO_CONJNOTRANS || transb < MFO_NOTRANS || transb > MFO_CONJNOTRANS)
        return -1;
    int kind = has_conj(transa) ? (has_conj(transb) ? 3 : 1) : (has_conj(transb) ? 2 : 0);
    int ta = remove_conj(transa), tb = remove_conj(transb);
    int alpha_one = alpha[0] == 1.0 && alpha[1] == 0.0;
    int beta_zero = beta[0] == 0.0 && beta[1] == 0.0;
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < m; i++) {
            const double* x;
            int incx;
            const double* y;
            int incy;
            if (ta == MFO_NOTRANS) { x = a + 2 * (size_t)i; incx = lda; }
            else { x = a + 2 * (size_t)i * lda; incx = 1; }
            if (tb == MFO_NOTRANS) { y = b + 2 * (size_t)j * ldb; incy = 1; }
            else { y = b + 2 * (size_t)j; incy = ldb; }
            double dr, di;
            zdot(kind, k, x, incx, y, incy, &dr, &di);
            double tr = dr, ti = di;
            if (!alpha_one) cmul(alpha[0], alpha[1], dr, di, &tr, &ti);
            double* cij = c + 2 * (i + (size_t)j * ldc);
            if (beta_zero) {
                cij[0] = tr;
                cij[1] = ti;
            } else {
                double ur, ui;
                cmul(beta[0], beta[1], cij[0], cij[1], &ur, &ui);
                cij[0] = tr + ur;
                cij[1] = ti + ui;
            }
        }
    }
    return 0;
}

void mfo_dposv_batch(int count, const int* n, const int* lda, const int64_t* off,
                     double* a, const int64_t* xoff, double* x, int* info, int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int i = 0; i < count; i++) {
        info[i] = mfo_dposv(n[i], a + off[i], lda[i], x + xoff[i]);
    }
    (void)threads;
}

void mfo_dgetrf_batch(int count, int m, int n, int lda, int64_t stride, double* a,
                      int* piv, int* sign, int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int i = 0; i < count; i++) {
        sign[i] = mfo_dgetrf(m, n, a + (size_t)i * stride, lda, piv + (size_t)i * m);
    }
    (void)threads;
}

/* LU solve over a fixed-size batch: per matrix mfo_dgetrf then mfo_dgetrs
 * (the LuInverse composition for one right-hand side), a overwritten by the
 * factors, piv[i*n..], x[i*n..] from b[i*n..]. */
void mfo_dgesv_batch(int count, int n, int lda, int64_t stride, double* a, int* piv,
                     const double* b, double* x, int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int i = 0; i < count; i++) {
        double* ai = a + (size_t)i * stride;
        int* pi = piv + (size_t)i * n;
        mfo_dgetrf(n, n, ai, lda, pi);
        mfo_dgetrs(n, ai, lda, pi, b + (size_t)i * n, x + (size_t)i * n);
    }
    (void)threads;
}

/* Products C_i = alpha_i op(A_i) op(B_i) + beta_i C_i over a batch (mfo_dgemm,
 * or mfo_zgemm when cplx): operands at a + offa[i], ... (offsets in doubles;
 * complex scalars and entries are (re, im) pairs). */
void mfo_gemm_vbatch(int cplx, int count, const int* ta, const int* tb, const int* m, const int* n,
                     const int* k, const double* alpha, const double* a, const int64_t* offa,
                     const int* lda, const double* b, const int64_t* offb, const int* ldb,
                     const double* beta, double* c, const int64_t* offc, const int* ldc, int* info,
                     int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
    for (int i = 0; i < count; i++) {
        if (cplx)
            info[i] = mfo_zgemm(ta[i], tb[i], m[i], n[i], k[i], alpha + 2 * (size_t)i, a + offa[i], lda[i],
                                b + offb[i], ldb[i], beta + 2 * (size_t)i, c + offc[i], ldc[i]);
        else
            info[i] = mfo_dgemm(ta[i], tb[i], m[i], n[i], k[i], alpha[i], a + offa[i], lda[i], b + offb[i],
                                ldb[i], beta[i], c + offc[i], ldc[i]);
    }
    (void)threads;
}

/* mfo_dposv_backward over a fixed-size batch: L, Abar stride elements
 * apart, x, w, bbar n apart. */
void mfo_dposv_backward_batch(int count, int n, int lda, int64_t stride, const double* l,
                              const double* x, const double* w, double* abar, double* bbar,
                              int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
    for (int i = 0; i < count; i++) {
        mfo_dposv_backward(n, l + (size_t)i * stride, lda, x + (size_t)i * n, w + (size_t)i * n,
                           abar + (size_t)i * stride, lda, bbar + (size_t)i * n);
    }
    (void)threads;
}

/* Adjoint of x = A^-1 b, A = L L^T (no reference counterpart; MatFlat has no
 * derivatives).  bbar = L^-T (L^-1 w) through the same two solves as
 * mfo_dposv; Abar = -(bbar x^T + x bbar^T)/2. */
void mfo_dposv_backward(int n, const double* l, int lda, const double* x, const double* w,
                        double* abar, int ldab, double* bbar)
{
    for (int i = 0; i < n; i++) bbar[i] = w[i];
    mfo_dtrsv(MFO_LOWER, MFO_NOTRANS, n, l, lda, bbar, 1);
    mfo_dtrsv(MFO_LOWER, MFO_TRANS, n, l, lda, bbar, 1);
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) {
            double u = bbar[i] * x[j];
            double v = x[i] * bbar[j];
            abar[i + (size_t)j * ldab] = -0.5 * (u + v);
        }
    }
}

/* Adjoint of the Cholesky factorization A = L L^T (no reference
 * counterpart; MatFlat has no derivatives).  Given L (lower, as
 * mfo_dpotrf leaves it) and Lbar = dloss/dL (lower triangle read), the
 * gradient with respect to the symmetric matrix A is
 *     Abar = L^-T M L^-1,  M = (Phi(L^T Lbar) + Phi(L^T Lbar)^T) / 2,
 * Phi = lower triangle with the diagonal halved (I. Murray, "Differentiation
 * of the Cholesky decomposition", arXiv:1602.07527, 2016, eq. 10, taken
 * through the symmetric part).  Operation order, which the GPU kernel
 * follows exactly:
 *   M[i][j] = M[j][i] = 0.5 * s, s = sum over k = i..n-1 of L[k][i] Lbar[k][j]
 *       (i >= j, s from 0.0 upward in k);
 *   every column x of M: L^T x = m solved in dot form, j = n-1 down to 0:
 *       x[j] = (x[j] - Dot(n-1-j, L[j+1..][j], x[j+1..])) / L[j][j] with
 *       Internals.Dot's pairing (Internals.cs:265-292) -> X = L^-T M;
 *   every row of X: the same solve along the row (stride ldab) -> X L^-1;
 *   Abar[i][j] = Abar[j][i] = 0.5 * (Abar[i][j] + Abar[j][i]) for i > j. */
static void lt_solve_dot(int n, const double* l, int ldl, double* x, int incx)
{
    for (int j = n - 1; j >= 0; j--) {
        const double d = mfo_ddot(n - 1 - j, l + (j + 1) + (size_t)j * ldl, 1, x + (size_t)(j + 1) * incx, incx);
        x[(size_t)j * incx] = (x[(size_t)j * incx] - d) / l[j + (size_t)j * ldl];
    }
}

void mfo_dpotrf_backward(int n, const double* l, int ldl, const double* lbar, int ldlb, double* abar, int ldab)
{
    for (int j = 0; j < n; j++) {
        for (int i = j; i < n; i++) {
            double s = 0.0;
            for (int k = i; k < n; k++) s = s + l[k + (size_t)i * ldl] * lbar[k + (size_t)j * ldlb];
            const double m = 0.5 * s;
            abar[i + (size_t)j * ldab] = m;
            abar[j + (size_t)i * ldab] = m;
        }
    }
    for (int c = 0; c < n; c++) lt_solve_dot(n, l, ldl, abar + (size_t)c * ldab, 1);
    for (int r = 0; r < n; r++) lt_solve_dot(n, l, ldl, abar + r, ldab);
    for (int j = 0; j < n; j++) {
        for (int i = j + 1; i < n; i++) {
            const double v = 0.5 * (abar[i + (size_t)j * ldab] + abar[j + (size_t)i * ldab]);
            abar[i + (size_t)j * ldab] = v;
            abar[j + (size_t)i * ldab] = v;
        }
    }
}

void mfo_dpotrf_backward_batch(int count, int n, int ld, int64_t stride, const double* l, const double* lbar,
                               double* abar, int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
    for (int i = 0; i < count; i++) {
        mfo_dpotrf_backward(n, l + (size_t)i * stride, ld, lbar + (size_t)i * stride, ld, abar + (size_t)i * stride,
                            ld);
    }
    (void)threads;
}

/* ------------------------------------------------------------------ complex
 * Complex128 factorizations and triangular solve.  Buffers hold interleaved
 * (re, im) doubles; element (i, j) of a matrix is at 2*(i + j*lda).
 *
 * System.Numerics.Complex operator / (BCL): Smith's method (R. L. Smith,
 * "Algorithm 116: Complex division", CACM 5(8):435, 1962) with the branch
 * taken on |d| < |c|:
 *   |d| <  |c|: r = d/c, den = c + d*r, q = ((a + b*r)/den, (b - a*r)/den)
 *   otherwise : r = c/d, den = d + c*r, q = ((b + a*r)/den, (-a + b*r)/den)
 * for (a + bi) / (c + di). */
void mfo_zdiv(double a, double b, double c, double d, double* re, double* im)
{
    if (fabs(d) < fabs(c)) {
        double r = d / c;
        double den = c + d * r;
        *re = (a + b * r) / den;
        *im = (b - a * r) / den;
    } else {
        double r = c / d;
        double den = d + c * r;
        *re = (b + a * r) / den;
        *im = (-a + b * r) / den;
    }
}

/* Internals.Dot(int n, Complex* x, int incx, Complex* y, int incy),
 * Internals.cs:294-321 (conj = 0), and Internals.DotConj(int n, Complex* x,
 * int incx, Complex* y, int incy), Internals.cs:352-380 (conj = 1, products
 * MulConj(x, y) = conj(x)*y).  Odd n: sum = p0 first; even n: sum = 0.  Then
 * pairs: sum = sum + (p_k + p_{k+1}), component-wise. */
static void zdot_pairs(int conj, int n, const double* x, int incx, const double* y, int incy,
                       double* sre, double* sim)
{
    double sr, si;
    if (n & 1) {
        if (conj) cmulconj(x, y, &sr, &si);
        else cmul(x[0], x[1], y[0], y[1], &sr, &si);
        x += 2 * (size_t)incx;
        y += 2 * (size_t)incy;
        n--;
    } else {
        sr = 0.0;
        si = 0.0;
    }
    while (n > 0) {
        double r0, i0, r1, i1;
        const double* x1 = x + 2 * (size_t)incx;
        const double* y1 = y + 2 * (size_t)incy;
        if (conj) {
            cmulconj(x, y, &r0, &i0);
            cmulconj(x1, y1, &r1, &i1);
        } else {
            cmul(x[0], x[1], y[0], y[1], &r0, &i0);
            cmul(x1[0], x1[1], y1[0], y1[1], &r1, &i1);
        }
        sr = sr + (r0 + r1);
        si = si + (i0 + i1);
        x += 4 * (size_t)incx;
        y += 4 * (size_t)incy;
        n -= 2;
    }
    *sre = sr;
    *sim = si;
}

void mfo_zdot(int conj, int n, const double* x, int incx, const double* y, int incy, double* out)
{
    zdot_pairs(conj, n, x, incx, y, incy, out, out + 1);
}

/* Factorization.Lu.cs:352-406 (LuCore, Complex).  As mfo_dgetrf, with the
 * pivot chosen by Internals.FastMagnitude = |re| + |im| (Internals.cs:15-19)
 * under the same strict '>' scan, the zero test diag != Complex.Zero, and
 * BCL division for the multipliers (Internals.DivInplace, Internals.cs:125-147). */
int mfo_zgetrf(int m, int n, double* a, int lda, int* piv)
{
    double* work = (double*)malloc(sizeof(double) * 2 * (size_t)m);
    for (int i = 0; i < m; i++) piv[i] = i;
    int sign = 1;
    for (int j = 0; j < n; j++) {
        double* colj = a + 2 * (size_t)j * lda;
        memcpy(work, colj, sizeof(double) * 2 * (size_t)m);
        for (int i = 0; i < m; i++) {
            double sr, si;
            zdot_pairs(0, imin(i, j), a + 2 * (size_t)i, lda, work, 1, &sr, &si);
            work[2 * i] = work[2 * i] - sr;
            work[2 * i + 1] = work[2 * i + 1] - si;
            colj[2 * i] = work[2 * i];
            colj[2 * i + 1] = work[2 * i + 1];
        }
        int p = j;
        for (int i = j + 1; i < m; i++) {
            double mi = fabs(work[2 * i]) + fabs(work[2 * i + 1]);
            double mp = fabs(work[2 * p]) + fabs(work[2 * p + 1]);
            if (mi > mp) p = i;
        }
        if (p != j) {
            for (int c = 0; c < n; c++) {
                double* x = a + 2 * (p + (size_t)c * lda);
                double* y = a + 2 * (j + (size_t)c * lda);
                double t0 = x[0], t1 = x[1];
                x[0] = y[0];
                x[1] = y[1];
                y[0] = t0;
                y[1] = t1;
            }
            int t = piv[p];
            piv[p] = piv[j];
            piv[j] = t;
            sign = -sign;
        }
        if (j < m) {
            double dr = colj[2 * j], di = colj[2 * j + 1];
            if (dr != 0.0 || di != 0.0) {
                for (int i = j + 1; i < m; i++) {
                    double qr, qi;
                    mfo_zdiv(colj[2 * i], colj[2 * i + 1], dr, di, &qr, &qi);
                    colj[2 * i] = qr;
                    colj[2 * i + 1] = qi;
                }
            }
        }
    }
    free(work);
    return sign;
}

/* Factorization.Cholesky.cs:195-249 (Complex, A = L L^H).  For row j and
 * k < j: t = (A[k,j] - DotConj(k, L[j,0:], L[k,0:])) / L[k,k] by BCL
 * division, L[j,k] = conj(t), s = s + (t.re*t.re + t.im*t.im); then
 * s = re(A[j,j]) - s and L[j,j] = (sqrt(s), 0) if s > 0, else failure after
 * the partial writes (returns j + 1).  The strict upper triangle is zeroed
 * at the end. */
int mfo_zpotrf(int n, double* a, int lda)
{
    for (int j = 0; j < n; j++) {
        double* colj = a + 2 * (size_t)j * lda;
        double s = 0.0;
        for (int k = 0; k < j; k++) {
            double* colk = a + 2 * (size_t)k * lda;
            double dr, di, tr, ti;
            zdot_pairs(1, k, a + 2 * (size_t)j, lda, a + 2 * (size_t)k, lda, &dr, &di);
            mfo_zdiv(colj[2 * k] - dr, colj[2 * k + 1] - di, colk[2 * k], colk[2 * k + 1], &tr, &ti);
            colk[2 * j] = tr;
            colk[2 * j + 1] = -ti;
            s = s + (tr * tr + ti * ti);
        }
        s = colj[2 * j] - s;
        if (s > 0) {
            colj[2 * j] = sqrt(s);
            colj[2 * j + 1] = 0.0;
        } else {
            return j + 1;
        }
    }
    for (int j = 1; j < n; j++) {
        double* colj = a + 2 * (size_t)j * lda;
        for (int i = 0; i < 2 * j; i++) colj[i] = 0.0;
    }
    return 0;
}

/* Blas.SolveTriangular.cs:280-395 (Complex) with MulSub (:397-421,
 * dst -= x*y) and MulSubConj (:423-445, dst -= MulConj(x, y)).  All four
 * transposes; the ConjNoTrans / ConjTrans forms use conj(A).  Division is
 * BCL division (mfo_zdiv). */
int mfo_ztrsv(int uplo, int trans, int n, const double* a, int lda, double* x, int incx)
{
    const int conj = (trans == MFO_CONJNOTRANS || trans == MFO_CONJTRANS);
    const int t = conj ? remove_conj(trans) : trans;
    if (trans < MFO_NOTRANS || trans > MFO_CONJNOTRANS) return -2;
    if (uplo != MFO_UPPER && uplo != MFO_LOWER) return -1;
    if ((uplo == MFO_UPPER) == (t == MFO_NOTRANS)) {
        /* axpy forms: Upper/NoTrans (:317-326, :337-346 conj) walks i down and
         * subtracts column i; Lower/Trans (:362-371, :381-390 conj) walks j down
         * and subtracts row j. */
        for (int i = n - 1; i >= 0; i--) {
            double* xp = x + 2 * (size_t)incx * i;
            const double* d = a + 2 * ((size_t)lda * i + i);
            double qr, qi;
            mfo_zdiv(xp[0], xp[1], d[0], conj ? -d[1] : d[1], &qr, &qi);
            xp[0] = qr;
            xp[1] = qi;
            for (int r = 0; r < i; r++) {
                const double* e = (uplo == MFO_UPPER) ? a + 2 * ((size_t)lda * i + r) : a + 2 * (i + (size_t)lda * r);
                double pr, pi;
                if (conj) cmulconj(e, xp, &pr, &pi);
                else cmul(e[0], e[1], xp[0], xp[1], &pr, &pi);
                double* xr = x + 2 * (size_t)incx * r;
                xr[0] = xr[0] - pr;
                xr[1] = xr[1] - pi;
            }
        }
    } else {
        /* dot forms: Upper/Trans (:327-334, :347-354 conj) dots column j of A
         * with x[0:j]; Lower/NoTrans (:355-361, :372-379 conj) dots row i. */
        for (int i = 0; i < n; i++) {
            double* xp = x + 2 * (size_t)incx * i;
            double sr, si;
            if (uplo == MFO_UPPER) zdot_pairs(conj, i, a + 2 * (size_t)lda * i, 1, x, incx, &sr, &si);
            else zdot_pairs(conj, i, a + 2 * (size_t)i, lda, x, incx, &sr, &si);
            const double* d = a + 2 * ((size_t)lda * i + i);
            double qr, qi;
            mfo_zdiv(xp[0] - sr, xp[1] - si, d[0], conj ? -d[1] : d[1], &qr, &qi);
            xp[0] = qr;
            xp[1] = qi;
        }
    }
    return 0;
}

/* ---- generalized symmetric-definite eigenproblem ------------------------ */

/* Internals.Norm(int n, double* x), Internals.cs:435-460: odd n starts with
 * x0*x0, even n with 0.0, then sum + (x_k*x_k + x_{k+1}*x_{k+1}); sqrt. */
double mfo_dnorm(int n, const double* x)
{
    double s;
    int c;
    if (n & 1) {
        s = x[0] * x[0];
        c = 1;
    } else {
        s = 0.0;
        c = 0;
    }
    for (; c < n; c += 2) s = s + (x[c] * x[c] + x[c + 1] * x[c + 1]);
    return sqrt(s);
}

/* Internals.MulAdd (Internals.cs:49-73): dst[i] += x[i]*y, element-wise. */
static void muladd(int n, const double* x, double y, double* dst)
{
    for (int i = 0; i < n; i++) dst[i] = dst[i] + x[i] * y;
}

/* Math.Max(double, double) of .NET: NaN if either argument is NaN. */
static double net_max(double a, double b)
{
    if (a != b) {
        if (!isnan(a)) return b < a ? a : b;
        return a;
    }
    return signbit(b) ? a : b;
}

/* Internals.Drotg (Internals.cs:626-668), the reference BLAS drotg. */
void mfo_drotg(double* da, double* db, double* c, double* s)
{
    double r, z;
    double roe = *db;
    double ada = fabs(*da);
    double adb = fabs(*db);
    if (ada > adb) roe = *da;
    double scale = ada + adb;
    if (scale == 0.0) {
        *c = 1.0;
        *s = 0.0;
        r = 0.0;
        z = 0.0;
    } else {
        double sda = *da / scale;
        double sdb = *db / scale;
        r = scale * sqrt(sda * sda + sdb * sdb);
        if (roe < 0.0) r = -r;
        *c = *da / r;
        *s = *db / r;
        z = 1.0;
        if (ada > adb) z = *s;
        if (adb >= ada && *c != 0.0) z = 1.0 / *c;
    }
    *da = r;
    *db = z;
}

/* Passes of the singular-value loop after which a block is reported as not
 * converged: 1000 QR steps (the reference's limit) plus, around each of
 * them, the at most n+1 deflation passes that can separate two steps.  A
 * finite matrix never reaches it; the reference loops without end on some
 * NaN patterns, which end here instead. */
static long svd_pass_cap(int n) { return 1001L * (n + 2) + 1; }

/* Factorization.Svd.cs:786-1256 (SvdCore double, LINPACK dsvdc order) for a
 * square n x n matrix with the left vectors only (vt == null, as
 * Factorization.Gevd.cs:366 calls it): a is destroyed, s gets the singular
 * values in descending order, u (ldu) the left singular vectors.  work, e
 * and stmp hold n doubles each.  Returns 0, or 1 when the iteration did not
 * converge (the reference throws MatrixFactorizationException; u then holds
 * the vectors reached so far and s is not written). */
int mfo_dsvd_u(int n, double* a, int lda, double* s, double* u, int ldu, double* work, double* e,
               double* stmp)
{
    const int m = n;
    const int nct = m - 1 < n ? m - 1 : n;
    int nrt = n - 2 < m ? n - 2 : m;
    if (nrt < 0) nrt = 0;
    const int kmax = nct > nrt ? nct : nrt;
    for (int k = 0; k < kmax; k++) {
        const int kp1 = k + 1;
        double* ak = a + (size_t)lda * k;
        if (k < nct) {
            stmp[k] = mfo_dnorm(m - k, ak + k);
            if (stmp[k] != 0.0) {
                if (ak[k] < 0.0) stmp[k] = -stmp[k];
                for (int i = k; i < m; i++) ak[i] = ak[i] / stmp[k];
                ak[k] = ak[k] + 1.0;
            }
            stmp[k] = -stmp[k];
        }
        for (int j = kp1; j < n; j++) {
            double* aj = a + (size_t)lda * j;
            if (k < nct && stmp[k] != 0.0) {
                double t = -mfo_ddot(m - k, ak + k, 1, aj + k, 1) / ak[k];
                muladd(m - k, ak + k, t, aj + k);
            }
            e[j] = aj[k];
        }
        if (k < nct) memcpy(u + (size_t)ldu * k + k, ak + k, sizeof(double) * (m - k));
        if (k < nrt) {
            e[k] = mfo_dnorm(n - kp1, e + kp1);
            if (e[k] != 0.0) {
                if (e[kp1] < 0.0) e[k] = -e[k];
                for (int i = kp1; i < n; i++) e[i] = e[i] / e[k];
                e[kp1] = e[kp1] + 1.0;
            }
            e[k] = -e[k];
            if (kp1 < m && e[k] != 0.0) {
                for (int i = kp1; i < m; i++) work[i] = 0.0;
                for (int j = kp1; j < n; j++) muladd(m - kp1, a + (size_t)lda * j + kp1, e[j], work + kp1);
                for (int j = kp1; j < n; j++) muladd(m - kp1, work + kp1, -e[j] / e[kp1], a + (size_t)lda * j + kp1);
            }
        }
    }

    int p = n < m + 1 ? n : m + 1;
    if (nct < n) stmp[nct] = a[(size_t)nct * lda + nct];
    if (m < p) stmp[p - 1] = 0.0;
    if (nrt + 1 < p) e[nrt] = a[(size_t)(p - 1) * lda + nrt];
    e[p - 1] = 0.0;

    for (int j = nct; j < m; j++) {
        double* uj = u + (size_t)ldu * j;
        for (int i = 0; i < m; i++) uj[i] = 0.0;
        uj[j] = 1.0;
    }
    for (int k = nct - 1; k >= 0; k--) {
        double* uk = u + (size_t)ldu * k;
        if (stmp[k] != 0.0) {
            for (int j = k + 1; j < m; j++) {
                double* uj = u + (size_t)ldu * j;
                double t = -mfo_ddot(m - k, uk + k, 1, uj + k, 1) / uk[k];
                muladd(m - k, uk + k, t, uj + k);
            }
            for (int i = k; i < m; i++) uk[i] = -uk[i];
            uk[k] = uk[k] + 1.0;
            for (int i = 0; i < k; i++) uk[i] = 0.0;
        } else {
            for (int i = 0; i < m; i++) uk[i] = 0.0;
            uk[k] = 1.0;
        }
    }

    const int pp = p - 1;
    int iter = 0;
    long passes = 0;
    const double eps = 0x1p-52;
    while (p > 0) {
        if (iter >= 1000 || ++passes > svd_pass_cap(n)) return 1;
        int k;
        for (k = p - 2; k >= 0; k--) {
            if (fabs(e[k]) <= eps * (fabs(stmp[k]) + fabs(stmp[k + 1]))) {
                e[k] = 0.0;
                break;
            }
        }
        int kase;
        if (k == p - 2) {
            kase = 4;
        } else {
            int ks;
            for (ks = p - 1; ks > k; ks--) {
                double t = 0.0;
                if (ks != p - 1) t = t + fabs(e[ks]);
                if (ks != k + 1) t = t + fabs(e[ks - 1]);
                if (fabs(stmp[ks]) <= eps * t) {
                    stmp[ks] = 0.0;
                    break;
                }
            }
            if (ks == k) {
                kase = 3;
            } else if (ks == p - 1) {
                kase = 1;
            } else {
                kase = 2;
                k = ks;
            }
        }
        k++;

        double f, cs, sn;
        if (kase == 1) {
            f = e[p - 2];
            e[p - 2] = 0.0;
            for (int j = k; j < p - 1; j++) {
                int l = p - 2 - j + k;
                double t = stmp[l];
                mfo_drotg(&t, &f, &cs, &sn);
                stmp[l] = t;
                if (l != k) {
                    f = -sn * e[l - 1];
                    e[l - 1] = e[l - 1] * cs;
                }
            }
        } else if (kase == 2) {
            f = e[k - 1];
            e[k - 1] = 0.0;
            for (int j = k; j < p; j++) {
                double t = stmp[j];
                mfo_drotg(&t, &f, &cs, &sn);
                stmp[j] = t;
                f = -sn * e[j];
                e[j] = cs * e[j];
                double* uj = u + (size_t)ldu * j;
                double* ukm1 = u + (size_t)ldu * (k - 1);
                for (int i = 0; i < m; i++) {
                    double z = cs * uj[i] + sn * ukm1[i];
                    ukm1[i] = cs * ukm1[i] - sn * uj[i];
                    uj[i] = z;
                }
            }
        } else if (kase == 3) {
            double scale = 0.0;
            scale = net_max(scale, fabs(stmp[p - 1]));
            scale = net_max(scale, fabs(stmp[p - 2]));
            scale = net_max(scale, fabs(e[p - 2]));
            scale = net_max(scale, fabs(stmp[k]));
            scale = net_max(scale, fabs(e[k]));
            double sp = stmp[p - 1] / scale;
            double spm1 = stmp[p - 2] / scale;
            double epm1 = e[p - 2] / scale;
            double sk = stmp[k] / scale;
            double ek = e[k] / scale;
            double b = ((spm1 + sp) * (spm1 - sp) + epm1 * epm1) / 2.0;
            double c = (sp * epm1) * (sp * epm1);
            double shift = 0.0;
            if (b != 0.0 || c != 0.0) {
                shift = sqrt(b * b + c);
                if (b < 0.0) shift = -shift;
                shift = c / (b + shift);
            }
            f = (sk + sp) * (sk - sp) + shift;
            double g = sk * ek;
            for (int j = k; j < p - 1; j++) {
                mfo_drotg(&f, &g, &cs, &sn);
                if (j != k) e[j - 1] = f;
                f = cs * stmp[j] + sn * e[j];
                e[j] = cs * e[j] - sn * stmp[j];
                g = sn * stmp[j + 1];
                stmp[j + 1] = stmp[j + 1] * cs;
                mfo_drotg(&f, &g, &cs, &sn);
                stmp[j] = f;
                f = cs * e[j] + sn * stmp[j + 1];
                stmp[j + 1] = -(sn * e[j]) + cs * stmp[j + 1];
                g = sn * e[j + 1];
                e[j + 1] = e[j + 1] * cs;
                if (j < m) {
                    double* uj = u + (size_t)ldu * j;
                    double* uj1 = u + (size_t)ldu * (j + 1);
                    for (int i = 0; i < m; i++) {
                        double z = cs * uj[i] + sn * uj1[i];
                        uj1[i] = cs * uj1[i] - sn * uj[i];
                        uj[i] = z;
                    }
                }
            }
            e[p - 2] = f;
            iter++;
        } else {
            if (stmp[k] < 0.0) stmp[k] = -stmp[k];
            while (k < pp) {
                if (stmp[k] >= stmp[k + 1]) break;
                double t = stmp[k];
                stmp[k] = stmp[k + 1];
                stmp[k + 1] = t;
                if (k < m) {
                    double* uk = u + (size_t)ldu * k;
                    double* uk1 = u + (size_t)ldu * (k + 1);
                    for (int i = 0; i < m; i++) {
                        double z = uk[i];
                        uk[i] = uk1[i];
                        uk1[i] = z;
                    }
                }
                k++;
            }
            iter = 0;
            passes = 0;
            p--;
        }
    }
    memcpy(s, stmp, sizeof(double) * n);
    return 0;
}

/* Factorization.Gevd double (Factorization.Gevd.cs:145-186, GevdCore
 * :330-376): B = L L^T in place (b), C = L^-1 A L^-T from A's upper
 * triangle by Lower/NoTrans solves of every row, then of every column,
 * C = U S V^T, w = S, and the eigenvectors L^-T U overwrite a.  c holds
 * n*n doubles and work 3n.  Returns 0; j+1 when B's Cholesky fails at row j
 * (the reference throws MatFlatException; b then holds the partial writes
 * and a is untouched); n+1 when the SVD did not converge (a holds U so far,
 * w is not written). */
int mfo_dgevd(int n, double* a, int lda, double* b, int ldb, double* w, double* c, double* work)
{
    int info = mfo_dpotrf(n, b, ldb);
    if (info != 0) return info;
    for (int j = 0; j < n; j++) {
        const double* aj = a + (size_t)lda * j;
        for (int i = 0; i < j; i++) {
            c[(size_t)n * j + i] = aj[i];
            c[(size_t)n * i + j] = aj[i];
        }
        c[(size_t)n * j + j] = aj[j];
    }
    for (int i = 0; i < n; i++) mfo_dtrsv(MFO_LOWER, MFO_NOTRANS, n, b, ldb, c + i, n);
    for (int i = 0; i < n; i++) mfo_dtrsv(MFO_LOWER, MFO_NOTRANS, n, b, ldb, c + (size_t)n * i, 1);
    if (mfo_dsvd_u(n, c, n, w, a, lda, work, work + n, work + 2 * n) != 0) return n + 1;
    for (int i = 0; i < n; i++) mfo_dtrsv(MFO_LOWER, MFO_TRANS, n, b, ldb, a + (size_t)lda * i, 1);
    return 0;
}

/* mfo_dgevd over a batch stored back-to-back (offsets in doubles). */
void mfo_dgevd_batch(int count, const int* n, const int* lda, const int* ldb, const int64_t* aoff,
                     const int64_t* boff, const int64_t* woff, double* a, double* b, double* w,
                     int* info, int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int i = 0; i < count; i++) {
        double* c = (double*)malloc(sizeof(double) * ((size_t)n[i] * n[i] + 3 * (size_t)n[i]));
        info[i] = c ? mfo_dgevd(n[i], a + aoff[i], lda[i], b + boff[i], ldb[i], w + woff[i], c,
                                c + (size_t)n[i] * n[i])
                    : -1;
        free(c);
    }
    (void)threads;
}
