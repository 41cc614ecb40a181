// mfh_general.hip -- MatFlat-order LU, Cholesky and triangular solve for
// complex128 matrices of every order and for double matrices above the
// LDS-resident range of mfh_exact.hip (129 .. MFH_EXACT_MAX_N rows).
//
// The arithmetic is written once over an element type: double, or zc for
// System.Numerics.Complex, whose operations restate the BCL's:
//  * a * b: re = a.re b.re - a.im b.im, im = a.im b.re + a.re b.im;
//  * a + b, a - b component-wise;
//  * a / b: Smith's method (R. L. Smith, CACM 5(8), 1962) with the branch
//    taken on |b.im| < |b.re|;
//  * Internals.MulConj(x, y) = conj(x) y (Internals.cs:21-29);
//  * Internals.FastMagnitude = |re| + |im| (Internals.cs:15-19), the LU
//    pivot magnitude.
// Dot products keep Internals.Dot's pairing (Internals.cs:265-321, and
// DotConj :352-380): odd length starts with the first product, even length
// with zero, then sum = sum + (p_k + p_{k+1}).  Every output element is
// produced by the same operations in the same order as MatFlat's
// sequential loops, so the results are bit-identical to them.
//
// Scheduling (MI355X): one workgroup of 256 threads (four wavefronts) per
// matrix, persistent over the batch with a fixed grid stride (no atomics).
// A matrix is staged in LDS when the launch's largest one fits the LDS
// budget; otherwise LU and the triangular solve work on the caller's matrix
// where it lies (its rows stay in L2 between the steps), and Cholesky works
// in a per-workgroup HBM scratch (its failure semantics forbid writing the
// caller's lower triangle before the failing row is known).  Both cases go
// through one generic pointer, so one kernel body serves both.
#include "mfh_device.h"
#include "mfh_internal.h"

#include <climits>

#pragma clang fp contract(off)

namespace {

constexpr int GT = 256;                 // threads per workgroup
constexpr int GR = MFH_EXACT_MAX_N / GT;  // rows per thread at the largest order
constexpr size_t kStageBytes = 64 * 1024;  // LDS for a staged matrix (two workgroups per CU)

struct zc {
    double re, im;
};

template <typename E>
struct El;

template <>
struct El<double> {
    __device__ static double zero() { return 0.0; }
    __device__ static double mul(double a, double b) { return a * b; }
    __device__ static double mulc(double a, double b) { return a * b; }
    __device__ static double add(double a, double b) { return a + b; }
    __device__ static double sub(double a, double b) { return a - b; }
    __device__ static double div(double a, double b) { return a / b; }
    __device__ static double conj(double a) { return a; }
    __device__ static bool nonzero(double a) { return a != 0.0; }
    __device__ static double mag(double a) { return __builtin_fabs(a); }
    __device__ static double re(double a) { return a; }
    __device__ static double from_re(double r) { return r; }
    __device__ static double norm2(double a) { return a * a; }
    __device__ static double nan() { return __builtin_nan(""); }
};

template <>
struct El<zc> {
    __device__ static zc zero() { return zc{0.0, 0.0}; }
    __device__ static zc mul(zc a, zc b) { return zc{a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im}; }
    // Internals.MulConj: a' = x.re, b' = -x.im, c = y.re, d = y.im;
    // (a'c - b'd, a'd + b'c)
    __device__ static zc mulc(zc x, zc y)
    {
        const double a = x.re, b = -x.im, c = y.re, d = y.im;
        return zc{a * c - b * d, a * d + b * c};
    }
    __device__ static zc add(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }
    __device__ static zc sub(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }
    __device__ static zc div(zc x, zc y)
    {
        const double a = x.re, b = x.im, c = y.re, d = y.im;
        if (__builtin_fabs(d) < __builtin_fabs(c)) {
            const double r = d / c;
            const double den = c + d * r;
            return zc{(a + b * r) / den, (b - a * r) / den};
        }
        const double r = c / d;
        const double den = d + c * r;
        return zc{(b + a * r) / den, (-a + b * r) / den};
    }
    __device__ static zc conj(zc a) { return zc{a.re, -a.im}; }
    __device__ static bool nonzero(zc a) { return a.re != 0.0 || a.im != 0.0; }
    __device__ static double mag(zc a) { return __builtin_fabs(a.re) + __builtin_fabs(a.im); }
    __device__ static double re(zc a) { return a.re; }
    __device__ static zc from_re(double r) { return zc{r, 0.0}; }
    __device__ static double norm2(zc a) { return a.re * a.re + a.im * a.im; }
    __device__ static zc nan() { return zc{__builtin_nan(""), __builtin_nan("")}; }
};

// c ? a : b component-wise (a select of a whole struct would go through
// memory)
__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }
__device__ __forceinline__ zc sel(bool c, zc a, zc b) { return zc{c ? a.re : b.re, c ? a.im : b.im}; }

template <typename E, bool CJ>
__device__ __forceinline__ E prod(E x, E y)
{
    return CJ ? El<E>::mulc(x, y) : El<E>::mul(x, y);
}

// Internals.Dot / DotConj in MatFlat's pairing; four products are issued
// per step (their loads and products do not depend on the sum) and only
// the additions into s stay sequential.
template <typename E, bool CJ>
__device__ __forceinline__ E dot_mf(int len, const E* x, int incx, const E* y, int incy)
{
    using O = El<E>;
    E s;
    int c;
    if (len & 1) {
        s = prod<E, CJ>(x[0], y[0]);
        c = 1;
    } else {
        s = O::zero();
        c = 0;
    }
    for (; c + 4 <= len; c += 4) {
        E p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) p[u] = prod<E, CJ>(x[(size_t)(c + u) * incx], y[(size_t)(c + u) * incy]);
        s = O::add(s, O::add(p[0], p[1]));
        s = O::add(s, O::add(p[2], p[3]));
    }
    for (; c < len; c += 2) {
        const E p0 = prod<E, CJ>(x[(size_t)c * incx], y[(size_t)c * incy]);
        const E p1 = prod<E, CJ>(x[(size_t)(c + 1) * incx], y[(size_t)(c + 1) * incy]);
        s = O::add(s, O::add(p0, p1));
    }
    return s;
}

// One product of a progressive dot-form substitution folded into row i's
// running Internals.Dot, in its pairing: length i odd -> s = p_0, then pairs
// (1,2),(3,4),...; even -> pairs (0,1),(2,3),...  c is uniform, the row
// parity is not, so the cases are selects.
template <typename E>
__device__ __forceinline__ void fold(int i, int c, E p, E& s, E& pend)
{
    const bool odd = i & 1;
    const bool codd = c & 1;
    const bool add = odd ? (c != 0 && !codd) : codd;
    const bool setp = odd ? codd : !codd;
    const bool sets = odd && c == 0;
    const E t = El<E>::add(s, El<E>::add(pend, p));
    s = sel(add, t, sel(sets, p, s));
    pend = sel(setp, p, pend);
}

__host__ __device__ inline size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

// ------------------------------------------------------------------- LU
template <typename E>
struct GenLu {
    const int* m;
    const int* n;
    E* const* a;
    const int* lda;
    int* const* piv;
    int m0, n0, lda0;
    E* a0;
    int64_t stride_a;
    int* piv0;
    int64_t stride_piv;
    int* sign;
    int* info;
    int max_m, max_n;
    int min_key;  // only matrices with max(m, n) > min_key (smaller ones have their own kernels)
    int report;   // write info for invalid matrices (else the size sort has done it)
    int stage;    // the matrix is staged in LDS (the launch's largest fits)
};

template <typename E>
__host__ __device__ size_t lu_fixed_bytes(int max_m, int max_n)
{
    return align16(align16((size_t)(max_m + max_n) * sizeof(E) + (size_t)max_m * sizeof(int)) + (GT / 64) * 12 + 16);
}

// Factorization.Lu.cs:296-350 (double) and :352-406 (Complex) in Crout
// order: at step k the candidates of column k (rows i >= k, dots of length k
// with the finished U column, staged in LDS) are formed in parallel, the
// pivot is a (magnitude, index) maximum with the lowest index on ties
// (MatFlat's strict '>' scan; a NaN magnitude ranks below every number, and
// a NaN at row k keeps p = k), rows k and p are swapped across all n
// columns, the multipliers are divided, and row k of U (columns j > k, dots
// of length k with the finished L row, staged in LDS) is formed in parallel.
template <typename E>
__global__ __launch_bounds__(GT) void k_getrf_gen(GenLu<E> B, int count)
{
    using O = El<E>;
    extern __shared__ __attribute__((aligned(16))) unsigned char glds[];
    E* colbuf = (E*)glds;
    E* rowbuf = colbuf + B.max_m;
    int* pv = (int*)(rowbuf + B.max_n);
    double* rkey = (double*)(glds + align16((size_t)((unsigned char*)(pv + B.max_m) - glds)));
    int* ridx = (int*)(rkey + GT / 64);
    int* sp = ridx + GT / 64;
    E* mat = (E*)(glds + lu_fixed_bytes<E>(B.max_m, B.max_n));
    const int tid = threadIdx.x;
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const int m = B.m ? B.m[b] : B.m0;
        const int n = B.n ? B.n[b] : B.n0;
        const int lda = B.lda ? B.lda[b] : B.lda0;
        E* A = B.a ? B.a[b] : B.a0 + (size_t)b * B.stride_a;
        int* P = B.piv ? B.piv[b] : B.piv0 + (size_t)b * B.stride_piv;
        int bad = 0;
        if (m < 1 || m > B.max_m || m > MFH_EXACT_MAX_N) bad = -3;
        else if (n < 1 || n > B.max_n || n > MFH_EXACT_MAX_N) bad = -4;
        else if (A == nullptr) bad = -5;
        else if (lda < m) bad = -6;
        else if (P == nullptr) bad = -7;
        if (bad) {
            if (B.report && tid == 0) B.info[b] = bad;
            continue;
        }
        if ((m > n ? m : n) <= B.min_key) continue;
        E* V = A;
        int ld = lda;
        if (B.stage) {
            V = mat;
            ld = m | 1;
            for (int e = tid; e < m * n; e += GT) {
                const int j = e / m, i = e - j * m;
                V[i + (size_t)j * ld] = A[i + (size_t)j * lda];
            }
        }
        for (int i = tid; i < m; i += GT) pv[i] = i;
        __syncthreads();
        int sign = 1;
        const int kmax = m < n ? m : n;
        for (int k = 0; k < kmax; ++k) {
            for (int i = tid; i < k; i += GT) colbuf[i] = V[i + (size_t)k * ld];
            __syncthreads();
            double key = -2.0;
            int idx = INT_MAX;
            bool wk_nan = false;
            for (int i = k + tid; i < m; i += GT) {
                const E w = O::sub(V[i + (size_t)k * ld], dot_mf<E, false>(k, V + i, ld, colbuf, 1));
                V[i + (size_t)k * ld] = w;
                const double mg = O::mag(w);
                const bool isn = mg != mg;
                if (i == k) wk_nan = isn;
                const double kk = isn ? -1.0 : mg;
                if (kk > key) {  // rows ascend within a thread: the first maximum stays
                    key = kk;
                    idx = i;
                }
            }
            mfh_wave_argmax(key, idx);
            if ((tid & 63) == 0) {
                rkey[tid >> 6] = key;
                ridx[tid >> 6] = idx;
            }
            __syncthreads();
            if (tid == 0) {  // thread 0 holds row k
                double kk = rkey[0];
                int ii = ridx[0];
                for (int w = 1; w < GT / 64; ++w) mfh_pivot_combine(kk, ii, rkey[w], ridx[w]);
                *sp = wk_nan ? k : ii;
            }
            __syncthreads();
            const int p = *sp;
            if (p != k) {
                for (int j = tid; j < n; j += GT) {
                    const E t = V[k + (size_t)j * ld];
                    V[k + (size_t)j * ld] = V[p + (size_t)j * ld];
                    V[p + (size_t)j * ld] = t;
                }
                if (tid == 0) {
                    const int t = pv[k];
                    pv[k] = pv[p];
                    pv[p] = t;
                }
                sign = -sign;
                __syncthreads();
            }
            const E diag = V[k + (size_t)k * ld];
            if (O::nonzero(diag))
                for (int i = k + 1 + tid; i < m; i += GT) V[i + (size_t)k * ld] = O::div(V[i + (size_t)k * ld], diag);
            for (int l = tid; l < k; l += GT) rowbuf[l] = V[k + (size_t)l * ld];
            __syncthreads();
            for (int j = k + 1 + tid; j < n; j += GT)
                V[k + (size_t)j * ld] = O::sub(V[k + (size_t)j * ld], dot_mf<E, false>(k, rowbuf, 1, V + (size_t)j * ld, 1));
            __syncthreads();
        }
        if (B.stage) {
            for (int e = tid; e < m * n; e += GT) {
                const int j = e / m, i = e - j * m;
                A[i + (size_t)j * lda] = V[i + (size_t)j * ld];
            }
        }
        for (int i = tid; i < m; i += GT) P[i] = pv[i];
        if (tid == 0) {
            B.sign[b] = sign;
            B.info[b] = 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- Cholesky
template <typename E>
struct GenChol {
    const int* n;
    E* const* a;
    const int* lda;
    int* info;
    int max_n;
    int min_key;
    int report;
    int stage;
    E* scratch;  // !stage: one max_n x max_n slot per workgroup (leading dimension max_n)
};

// Factorization.Cholesky.cs:114-168 (double) and :195-249 (Complex), by
// columns: the upper triangle is first transposed into M's lower triangle,
// so M[j, k] holds A[k, j] until step k replaces it with L[j, k].  At step
// k the owner of row k finishes s_k (its running sum of t*t, in MatFlat's
// order) and writes L[k, k] = sqrt(re(A[k, k]) - s_k), or records the
// failure; then every row j > k forms t = (A[k, j] - DotConj(k, L[j, :],
// L[k, :])) / L[k, k], stores L[j, k] = conj(t) and adds t*t (re^2 + im^2)
// to s_j -- exactly the operations the reference performs for row j at
// inner index k.  The caller's matrix is written only at the end: the whole
// factor with a zeroed strict upper triangle, or, after a failure at row f,
// exactly what the reference leaves when it throws (rows < f with their
// diagonal, row f without it, nothing else).
template <typename E>
__global__ __launch_bounds__(GT) void k_potrf_gen(GenChol<E> B, int count)
{
    using O = El<E>;
    extern __shared__ __attribute__((aligned(16))) unsigned char glds[];
    int* sfail = (int*)glds;
    E* mat = (E*)(glds + 16);
    const int tid = threadIdx.x;
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const int n = B.n[b];
        E* A = B.a[b];
        const int lda = B.lda[b];
        int bad = 0;
        if (n < 1 || n > B.max_n || n > MFH_EXACT_MAX_N) bad = -3;
        else if (A == nullptr) bad = -4;
        else if (lda < n || (long long)lda * n > MFH_MAX_MATRIX_ELEMS) bad = -5;
        if (bad) {
            if (B.report && tid == 0) B.info[b] = bad;
            continue;
        }
        if (n <= B.min_key) continue;
        E* M;
        int ld;
        if (B.stage) {
            M = mat;
            ld = n | 1;
        } else {
            M = B.scratch + (size_t)blockIdx.x * B.max_n * B.max_n;
            ld = B.max_n;
        }
        for (int e = tid; e < n * n; e += GT) {
            const int j = e / n, k = e - j * n;
            if (k <= j) M[j + (size_t)k * ld] = A[k + (size_t)j * lda];
        }
        if (tid == 0) *sfail = -1;
        double s[GR];
#pragma unroll
        for (int r = 0; r < GR; ++r) s[r] = 0.0;
        __syncthreads();
        int fail = -1;
        for (int k = 0; k < n; ++k) {
#pragma unroll
            for (int r = 0; r < GR; ++r) {
                if (tid + r * GT == k) {
                    const double sk = O::re(M[k + (size_t)k * ld]) - s[r];
                    if (sk > 0) M[k + (size_t)k * ld] = O::from_re(__builtin_sqrt(sk));
                    else *sfail = k;
                }
            }
            __syncthreads();
            // A later row's failure can be recorded before a slow wavefront
            // reads this step's flag; it carries that row's index, never k.
            if (*sfail == k) {
                fail = k;
                break;
            }
            const E dk = M[k + (size_t)k * ld];
#pragma unroll
            for (int r = 0; r < GR; ++r) {
                const int j = tid + r * GT;
                if (j > k && j < n) {
                    const E num = O::sub(M[j + (size_t)k * ld], dot_mf<E, true>(k, M + j, ld, M + k, ld));
                    const E t = O::div(num, dk);
                    M[j + (size_t)k * ld] = O::conj(t);
                    s[r] = s[r] + O::norm2(t);
                }
            }
        }
        if (fail < 0) {
            for (int e = tid; e < n * n; e += GT) {
                const int j = e / n, i = e - j * n;
                A[i + (size_t)j * lda] = (i >= j) ? M[i + (size_t)j * ld] : O::zero();
            }
            if (tid == 0) B.info[b] = 0;
        } else {
            for (int e = tid; e < n * n; e += GT) {
                const int j = e / n, i = e - j * n;
                if (i >= j && i <= fail && !(i == fail && j == fail)) A[i + (size_t)j * lda] = M[i + (size_t)j * ld];
            }
            if (tid == 0) B.info[b] = fail + 1;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------- triangular solve
template <typename E>
struct GenTrsv {
    const int* n;
    const E* const* a;
    const int* lda;
    E* const* x;
    const int* incx;  // null: incx0 for every matrix
    int incx0;
    int* info;
    int uplo, trans, unit;
    int max_n;
    int min_key;
    int report;
    int gate;      // composed solves: skip matrices whose info is not 0 and leave info alone
    int pos_n;     // argument position of n in the entry point (per-matrix info codes follow it)
};

// Blas.SolveTriangular.cs:162-242 (double) and :280-395 (Complex).  x lives
// in LDS, each element updated by its owner thread; A is read where it lies
// (every element is used once).
//  * Axpy forms -- Upper/NoTrans, Upper/ConjNoTrans, Lower/Trans,
//    Lower/ConjTrans: for i = n-1 .. 0 the owner divides x_i by the diagonal
//    (conjugated for the Conj forms), then every row r < i subtracts
//    e(r, i) x_i (MulSub :397-421) or conj(e(r, i)) x_i (MulSubConj :423-445),
//    in the reference's order for each x_r.
//  * Dot forms -- Upper/Trans, Upper/ConjTrans, Lower/NoTrans,
//    Lower/ConjNoTrans: x_i = (x_i - Dot(i, e(i, :), x)) / diag; each row
//    folds its products into its running Dot (or DotConj) as soon as the
//    x_c it needs is final, in Internals.Dot's pairing.
// diag MFH_UNIT (double only) skips the division: the unit-lower solve of
// Factorization.LuInverse.cs:174-182 divides by the 1.0 it stores, which is
// exact.
template <typename E>
__global__ __launch_bounds__(GT) void k_trsv_gen(GenTrsv<E> B, int count)
{
    using O = El<E>;
    extern __shared__ __attribute__((aligned(16))) unsigned char glds[];
    E* xs = (E*)glds;
    const int tid = threadIdx.x;
    const bool cj = B.trans == MFH_CONJNOTRANS || B.trans == MFH_CONJTRANS;
    const bool notrans = B.trans == MFH_NOTRANS || B.trans == MFH_CONJNOTRANS;
    const bool upper = B.uplo == MFH_UPPER;
    const bool unit = B.unit != 0;
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const int n = B.n[b];
        const E* A = B.a[b];
        const int lda = B.lda[b];
        E* X = B.x[b];
        const int incx = B.incx ? B.incx[b] : B.incx0;
        int bad = 0;
        if (n < 1 || n > B.max_n || n > MFH_EXACT_MAX_N) bad = -B.pos_n;
        else if (A == nullptr) bad = -(B.pos_n + 1);
        else if (lda < n) bad = -(B.pos_n + 2);
        else if (X == nullptr) bad = -(B.pos_n + 3);
        else if (incx < 1) bad = -(B.pos_n + 4);
        if (bad) {
            if (B.report && tid == 0) B.info[b] = bad;
            continue;
        }
        if (n <= B.min_key) continue;
        if (B.gate && B.info[b] != 0) continue;
        for (int i = tid; i < n; i += GT) xs[i] = X[(size_t)i * incx];
        __syncthreads();
        auto diag = [&](int i) {
            const E d = A[i + (size_t)i * lda];
            return sel(cj, O::conj(d), d);
        };
        if (upper == notrans) {
            // e(r, i): column i of U (Upper), row i of L (Lower, Trans)
            for (int i = n - 1; i >= 0; --i) {
                if (tid == (i & (GT - 1)) && !unit) xs[i] = O::div(xs[i], diag(i));
                __syncthreads();
                const E xi = xs[i];
#pragma unroll
                for (int r = 0; r < GR; ++r) {
                    const int rr = tid + r * GT;
                    if (rr < i) {
                        const E e = upper ? A[rr + (size_t)i * lda] : A[i + (size_t)rr * lda];
                        xs[rr] = O::sub(xs[rr], sel(cj, O::mulc(e, xi), O::mul(e, xi)));
                    }
                }
            }
        } else {
            // e(i, c): column i of U (Upper, Trans), row i of L (Lower)
            E s[GR], pend[GR];
#pragma unroll
            for (int r = 0; r < GR; ++r) {
                s[r] = O::zero();
                pend[r] = O::zero();
            }
            if (tid == 0) xs[0] = unit ? O::sub(xs[0], O::zero()) : O::div(O::sub(xs[0], O::zero()), diag(0));
            __syncthreads();
            for (int c = 0; c < n - 1; ++c) {
                const E xc = xs[c];
#pragma unroll
                for (int r = 0; r < GR; ++r) {
                    const int i = tid + r * GT;
                    if (i > c && i < n) {
                        const E e = upper ? A[c + (size_t)i * lda] : A[i + (size_t)c * lda];
                        fold(i, c, sel(cj, O::mulc(e, xc), O::mul(e, xc)), s[r], pend[r]);
                        if (i == c + 1) {
                            const E v = O::sub(xs[i], s[r]);
                            xs[i] = unit ? v : O::div(v, diag(i));
                        }
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
        for (int i = tid; i < n; i += GT) X[(size_t)i * incx] = xs[i];
        if (!B.gate && tid == 0) B.info[b] = 0;
        __syncthreads();
    }
}

// ------------------------------------------------------------ LU solve
// First step of the LU-solve composition (Factorization.LuInverse.cs:169-172
// for one right-hand side): x[j] = b[piv[j]], with b staged in LDS first so
// that b and x may alias.  An out-of-range pivot entry gives NaN.  The two
// triangular sweeps follow as k_trsv_gen launches gated on info.
struct GenPermute {
    MfhGetrsBatch B;
    int max_n;
    int min_key;
};

__global__ __launch_bounds__(GT) void k_getrs_permute(GenPermute G, int count)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char glds[];
    double* bs = (double*)glds;
    const int tid = threadIdx.x;
    const MfhGetrsBatch& B = G.B;
    const int max_n = G.max_n, min_key = G.min_key;
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const int n = B.n[b];
        int bad = 0;
        if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -3;
        else if (B.lu[b] == nullptr) bad = -4;
        else if (B.lda[b] < n) bad = -5;
        else if (B.piv[b] == nullptr) bad = -6;
        else if (B.b[b] == nullptr) bad = -7;
        else if (B.x[b] == nullptr) bad = -8;
        if (bad || n <= min_key) continue;  // the size sort reports invalid matrices
        const double* bv = B.b[b];
        const int* P = B.piv[b];
        double* X = B.x[b];
        for (int i = tid; i < n; i += GT) bs[i] = bv[i];
        __syncthreads();
        for (int i = tid; i < n; i += GT) {
            const int pr = P[i];
            X[i] = (pr >= 0 && pr < n) ? bs[pr] : __builtin_nan("");
        }
        if (tid == 0) B.info[b] = 0;
        __syncthreads();
    }
}

// ------------------------------------------------------------ launches
template <typename K, typename Arg>
int launch_gen(mfh_handle* h, int kind, int rows, K kernel, size_t lds, long max_groups, long cap, const Arg& arg,
               int count)
{
    MFH_HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, GT, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > cap) grid = cap;
    if (grid > max_groups) grid = max_groups;
    if (grid < 1) grid = 1;
    int st = mfh_timing_begin(h, kind, rows);
    if (st) return st;
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(GT), lds, h->stream, arg, count);
    MFH_HIP_CHECK(hipGetLastError());
    return mfh_timing_end(h);
}

template <typename E>
int launch_getrf(mfh_handle* h, int count, GenLu<E> B)
{
    const size_t fixed = lu_fixed_bytes<E>(B.max_m, B.max_n);
    const size_t mat = (size_t)(B.max_m | 1) * B.max_n * sizeof(E);
    B.stage = mat <= kStageBytes ? 1 : 0;
    const size_t lds = fixed + (B.stage ? mat : 0);
    const int rows = B.max_m > B.max_n ? B.max_m : B.max_n;
    return launch_gen(h, MFH_KERNEL_GETRF_GEN, rows, k_getrf_gen<E>, lds, count, LONG_MAX, B, count);
}

// The Cholesky scratch (when a matrix does not fit in LDS) is one slot per
// workgroup in the handle's second buffer; the grid is capped at two
// workgroups per CU so that the scratch stays at 2 * CUs * max_n^2 elements.
template <typename E>
int launch_potrf(mfh_handle* h, int count, GenChol<E> B)
{
    const size_t mat = (size_t)(B.max_n | 1) * B.max_n * sizeof(E);
    B.stage = mat <= kStageBytes ? 1 : 0;
    long cap = LONG_MAX;
    if (!B.stage) {
        cap = 2L * h->num_cus;
        if (cap > count) cap = count;
        int st = mfh_aux_reserve(h, (size_t)cap * B.max_n * B.max_n * sizeof(E));
        if (st) return st;
        B.scratch = (E*)h->aux;
    }
    const size_t lds = 16 + (B.stage ? mat : 0);
    return launch_gen(h, MFH_KERNEL_POTRF_GEN, B.max_n, k_potrf_gen<E>, lds, count, cap, B, count);
}

template <typename E>
int launch_trsv(mfh_handle* h, int count, const GenTrsv<E>& B)
{
    const size_t lds = align16((size_t)B.max_n * sizeof(E));
    return launch_gen(h, MFH_KERNEL_TRSV_GEN, B.max_n, k_trsv_gen<E>, lds, count, LONG_MAX, B, count);
}

}  // namespace

// Bytes of the handle's second buffer a Cholesky launch of this shape uses
// (its HBM scratch; 0 when every matrix is staged in LDS): plans reserve it
// before capture.
size_t mfh_potrf_gen_aux_bytes(const mfh_handle* h, int count, int max_n, bool cplx)
{
    if (!cplx && max_n <= MFH_LDS_MAX_N) return 0;
    const size_t elem = cplx ? sizeof(zc) : sizeof(double);
    if ((size_t)(max_n | 1) * max_n * elem <= kStageBytes) return 0;
    long cap = 2L * h->num_cus;
    if (cap > count) cap = count;
    return (size_t)cap * max_n * max_n * elem;
}

// ---------------------------------------------------------------- double
// Orders above the LDS-resident kernels' range (MFH_LDS_MAX_N); the size
// sort of mfh_exact.hip has reported the invalid matrices already.
int mfh_launch_getrf_gen_d(mfh_handle* h, int count, const MfhGetrfBatch& B, int max_m, int max_n)
{
    GenLu<double> G{B.m, B.n, B.a, B.lda, B.piv, B.m0, B.n0, B.lda0, B.a0, B.stride_a, B.piv0, B.stride_piv,
                    B.sign, B.info, max_m, max_n, MFH_LDS_MAX_N, 0, 0};
    return launch_getrf<double>(h, count, G);
}

int mfh_launch_potrf_gen_d(mfh_handle* h, int count, const MfhCholBatch& B, int max_n, bool posv)
{
    GenChol<double> G{B.n, B.a, B.lda, B.info, max_n, MFH_LDS_MAX_N, 0, 0, nullptr};
    int st = launch_potrf<double>(h, count, G);
    if (st || !posv) return st;
    // SolveTriangular(Lower, NoTrans) then (Lower, Trans) on x, for the
    // matrices whose factorization succeeded
    GenTrsv<double> T{B.n, (const double* const*)B.a, B.lda, B.x, nullptr, 1, B.info, MFH_LOWER, MFH_NOTRANS, 0,
                      max_n, MFH_LDS_MAX_N, 0, 1, 3};
    st = launch_trsv<double>(h, count, T);
    if (st) return st;
    T.trans = MFH_TRANS;
    return launch_trsv<double>(h, count, T);
}

int mfh_launch_trsv_gen_d(mfh_handle* h, int count, const MfhTrsvBatch& B, int max_n)
{
    GenTrsv<double> G{B.n, B.a, B.lda, B.x, B.incx, 1, B.info, B.uplo, B.trans, B.unit, max_n, MFH_LDS_MAX_N, 0, 0, 6};
    return launch_trsv<double>(h, count, G);
}

int mfh_launch_getrs_gen_d(mfh_handle* h, int count, const MfhGetrsBatch& B, int max_n)
{
    const size_t lds = align16((size_t)max_n * sizeof(double));
    const GenPermute P{B, max_n, MFH_LDS_MAX_N};
    int st = launch_gen(h, MFH_KERNEL_GETRS_PERMUTE, max_n, k_getrs_permute, lds, count, LONG_MAX, P, count);
    if (st) return st;
    // the unit-lower sweep, then Upper/NoTrans, on x in place
    GenTrsv<double> T{B.n, B.lu, B.lda, B.x, nullptr, 1, B.info, MFH_LOWER, MFH_NOTRANS, 1, max_n, MFH_LDS_MAX_N, 0,
                      1, 3};
    st = launch_trsv<double>(h, count, T);
    if (st) return st;
    T.uplo = MFH_UPPER;
    T.unit = 0;
    return launch_trsv<double>(h, count, T);
}

// --------------------------------------------------------------- complex
int mfh_launch_zgetrf(mfh_handle* h, int count, const int* m, const int* n, double* const* a, const int* lda,
                      int* const* piv, int* sign, int* info, int max_m, int max_n)
{
    GenLu<zc> G{m, n, (zc* const*)a, lda, piv, 0, 0, 0, nullptr, 0, nullptr, 0, sign, info, max_m, max_n, 0, 1, 0};
    return launch_getrf<zc>(h, count, G);
}

int mfh_launch_zpotrf(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, int* info, int max_n)
{
    GenChol<zc> G{n, (zc* const*)a, lda, info, max_n, 0, 1, 0, nullptr};
    return launch_potrf<zc>(h, count, G);
}

int mfh_launch_ztrsv(mfh_handle* h, int uplo, int trans, int count, const int* n, const double* const* a,
                     const int* lda, double* const* x, const int* incx, int* info, int max_n)
{
    GenTrsv<zc> G{n, (const zc* const*)a, lda, (zc* const*)x, incx, 1, info, uplo, trans, 0, max_n, 0, 1, 0, 5};
    return launch_trsv<zc>(h, count, G);
}
