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
