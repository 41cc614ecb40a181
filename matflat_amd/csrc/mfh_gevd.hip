// mfh_gevd.hip -- Factorization.Gevd (double): A v = w B v for symmetric A
// and symmetric positive definite B, in MatFlat's operation order.
//
// Factorization.Gevd.cs:330-376 (GevdCore) composes
//   B = L L^T                      Factorization.Cholesky (b is overwritten)
//   C = L^-1 A L^-T                Lower/NoTrans solves of every row of the
//                                  symmetric copy of A, then of every column
//   C = U S V^T                    SvdCore with left vectors only
//                                  (Factorization.Svd.cs:786-1256, LINPACK
//                                  dsvdc: Householder bidiagonalization, then
//                                  implicit-shift QR sweeps of Givens
//                                  rotations)
//   w = S, V = L^-T U              Lower/Trans solves of every column, in a.
//
// The Cholesky is the exact Cholesky kernel of this library, launched first.
// Everything after it runs in one wavefront per matrix (k_gevd_exact): C
// lives in a per-workgroup HBM matrix, U is built in a (as the reference
// builds it), s and e live in LDS, and
//   - every triangular solve, Householder dot and column update is one lane's
//     own sequential loop over a row or column, in the reference's order;
//   - the norms, the Householder scalars and the whole singular-value
//     iteration (the s and e recurrences, Drotg) are computed identically by
//     every lane, so they need no broadcast: all lanes write the same values
//     to the same LDS words;
//   - the Givens rotations and column swaps of U act within rows, so each
//     lane applies them to the rows it owns as the iteration produces them.
// Every value is therefore the same sequence of IEEE binary64 operations as
// the reference's, and the results are bit-identical to MatFlat's.
#include "mfh_device.h"
#include "mfh_internal.h"
#include "mfh_order.h"

#pragma clang fp contract(off)

namespace {

struct GevdBatch {
    const int* n;
    double* const* a;
    const int* lda;
    double* const* b;
    const int* ldb;
    double* const* w;
    int* info;
    int* n_chol;     // n where the matrix's arguments are valid, else 0 (the Cholesky skips it)
    int* info_chol;  // the Cholesky's info
    const int* order;  // problem ids, largest order first
    int* counter;      // work counter over order (zero before the launch)
    double* cws;       // C in HBM: max_n x ldc doubles per workgroup (when C is not in LDS)
};

// C lives in a per-workgroup HBM matrix (it stays in L2 while its problem
// runs) rather than in LDS: with only s and e in LDS, occupancy is set by
// registers, six problems per SIMD instead of one (10,000 pairs, n ~
// U{8..64}: 16.1 against 19.6 ms with C in LDS).  MFH_GEVD_C_IN_LDS=1 builds
// the LDS variant for comparison.
#ifndef MFH_GEVD_C_IN_LDS
#define MFH_GEVD_C_IN_LDS 0
#endif

constexpr int GW = 64;  // one wavefront per matrix
// the QR sweep carries at most two rows of U per lane
static_assert(MFH_SMALL_MAX_N <= 2 * GW, "rows of U per lane");

// Math.Max(double, double) of .NET: NaN when either argument is NaN.
__device__ __forceinline__ double net_max(double a, double b)
{
    if (a != b) {
        if (!__builtin_isnan(a)) return b < a ? a : b;
        return a;
    }
    return __builtin_signbit(b) ? a : b;
}

// Internals.Drotg, Internals.cs:626-668.
__device__ __forceinline__ void drotg(double& da, double& db, double& c, double& s)
{
    double r, z;
    double roe = db;
    const double ada = __builtin_fabs(da);
    const double adb = __builtin_fabs(db);
    if (ada > adb) roe = da;
    const double scale = ada + adb;
    if (scale == 0.0) {
        c = 1.0;
        s = 0.0;
        r = 0.0;
        z = 0.0;
    } else {
        const double sda = da / scale;
        const double sdb = db / scale;
        r = scale * __builtin_sqrt(sda * sda + sdb * sdb);
        if (roe < 0.0) r = -r;
        c = da / r;
        s = db / r;
        z = 1.0;
        if (ada > adb) z = s;
        if (adb >= ada && c != 0.0) z = 1.0 / c;
    }
    da = r;
    db = z;
}

// Internals.Norm(int n, double* x), Internals.cs:435-460.
__device__ __forceinline__ double norm_seq(int n, const double* x)
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
    return __builtin_sqrt(s);
}

// Passes of the singular-value loop after which a block is reported as not
// converged: the reference's 1000 QR steps, each with at most n + 1
// deflation passes around it.  A finite matrix never reaches it; the
// reference loops without end on some NaN patterns, which end here.
__device__ __forceinline__ long svd_pass_cap(int n) { return 1001L * (n + 2) + 1; }

// Rotation of the columns (c1, c2) of U within the rows this lane owns:
// z = cs u1 + sn u2; u2 = cs u2 - sn u1; u1 = z (Factorization.Svd.cs:1088-1093).
__device__ __forceinline__ void rot_rows(double* U, int ldu, int n, int lane, int c1, int c2, double cs, double sn)
{
    for (int i = lane; i < n; i += GW) {
        const double x = U[i + (size_t)c1 * ldu];
        const double y = U[i + (size_t)c2 * ldu];
        U[i + (size_t)c2 * ldu] = cs * y - sn * x;
        U[i + (size_t)c1 * ldu] = cs * x + sn * y;
    }
}

__global__ __launch_bounds__(GW) void k_gevd_check(GevdBatch G, int count, int max_n)
{
    const int i = blockIdx.x * GW + threadIdx.x;
    if (i >= count) return;
    const int n = G.n[i];
    int bad = 0;
    if (n < 1 || n > max_n || n > MFH_SMALL_MAX_N) bad = -3;
    else if (G.a[i] == nullptr) bad = -4;
    else if (G.lda[i] < n || (long long)G.lda[i] * n > MFH_MAX_MATRIX_ELEMS) bad = -5;
    else if (G.b[i] == nullptr) bad = -6;
    else if (G.ldb[i] < n || (long long)G.ldb[i] * n > MFH_MAX_MATRIX_ELEMS) bad = -7;
    else if (G.w[i] == nullptr) bad = -8;
    G.info[i] = bad;
    G.n_chol[i] = bad ? 0 : n;
}

// One wavefront per matrix; workgroups walk the batch with a fixed stride.
// LDS: C (n x ldc), U (n x ldc) when ULDS, then s and e (max_n each).
template <bool ULDS>
__global__ __launch_bounds__(GW) void k_gevd_exact(GevdBatch G, int count, int max_n, int ldc)
{
    extern __shared__ double sm[];
    constexpr int CL = MFH_GEVD_C_IN_LDS ? 1 : 0;  // matrices of LDS before S and E: C
    double* C = CL ? sm : G.cws + (size_t)blockIdx.x * max_n * ldc;
    double* sU = sm + (size_t)CL * max_n * ldc;
    double* S = sm + (size_t)((ULDS ? 1 : 0) + CL) * max_n * ldc;
    double* E = S + max_n;
    const int lane = threadIdx.x;

    // Problems are taken largest first through the work counter: with orders
    // mixed from 8 to 128 and the time of one growing as n^3, a fixed
    // stride over the batch left some workgroups with far more work than
    // others at the end.  Every test that ends or skips an iteration is on
    // a wave-uniform value (lane 0's slot, values read through
    // readfirstlane), and no path that skips an iteration holds a per-lane
    // branch: either makes the loop divergent to the compiler, which may
    // then re-run the top of the loop for part of the wavefront, where the
    // slot of the lanes other than lane 0 reads 0.
    for (;;) {
        int k = 0;
        if (lane == 0) k = atomicAdd(G.counter, 1);
        k = __builtin_amdgcn_readlane(k, 0);
        if (k >= count) break;
        const int mi = __builtin_amdgcn_readfirstlane(G.order[k]);
        if (__builtin_amdgcn_readfirstlane(G.info[mi]) != 0) continue;  // invalid arguments: nothing is written
        const int fail = __builtin_amdgcn_readfirstlane(G.info_chol[mi]);
        if (fail != 0) {  // B is not positive definite: b holds the Cholesky's partial writes
            G.info[mi] = fail;  // every lane stores the same value
            continue;
        }
        const int n = __builtin_amdgcn_readfirstlane(G.n[mi]);
        double* A = G.a[mi];
        const int lda = __builtin_amdgcn_readfirstlane(G.lda[mi]);
        const double* L = G.b[mi];
        const int ldb = __builtin_amdgcn_readfirstlane(G.ldb[mi]);
        double* U = ULDS ? sU : A;
        const int ldu = ULDS ? ldc : lda;
        __syncthreads();  // the previous matrix's last reads of LDS are done

        // C = the symmetric matrix whose upper triangle is A's
        // (Factorization.Gevd.cs:348-360); A's lower triangle is not read.
        for (int e = lane; e < n * n; e += GW) {
            const int j = e / n;
            const int i = e - j * n;
            C[i + j * ldc] = i <= j ? A[i + (size_t)j * lda] : A[j + (size_t)i * lda];
        }
        __syncthreads();
        // Lower/NoTrans solve of every row (:361-364): row r is a vector of stride ldc.
        for (int r = lane; r < n; r += GW) {
            for (int i = 0; i < n; ++i) {
                const double d = mfh_dot_exact(i, L + i, ldb, C + r, ldc);
                C[r + i * ldc] = (C[r + i * ldc] - d) / L[i + (size_t)i * ldb];
            }
        }
        __syncthreads();
        // ... then of every column (:365-368).
        for (int c = lane; c < n; c += GW) {
            double* x = C + c * ldc;
            for (int i = 0; i < n; ++i) {
                const double d = mfh_dot_exact(i, L + i, ldb, x, 1);
                x[i] = (x[i] - d) / L[i + (size_t)i * ldb];
            }
        }
        __syncthreads();

        // ---- SvdCore, square, left vectors only (Factorization.Svd.cs:786-1256)
        // Bidiagonalization: nct = n - 1 column and nrt = n - 2 row reflections.
        for (int k = 0; k < n - 1; ++k) {
            const int kp1 = k + 1;
            double* ck = C + k * ldc;
            double sk = norm_seq(n - k, ck + k);
            if (sk != 0.0) {
                if (ck[k] < 0.0) sk = -sk;
                for (int i = k + lane; i < n; i += GW) {
                    double v = ck[i] / sk;
                    if (i == k) v = v + 1.0;
                    ck[i] = v;
                    U[i + (size_t)k * ldu] = v;
                }
            } else {
                for (int i = k + lane; i < n; i += GW) U[i + (size_t)k * ldu] = ck[i];
            }
            S[k] = -sk;
            __syncthreads();
            for (int j = kp1 + lane; j < n; j += GW) {
                double* cj = C + j * ldc;
                if (sk != 0.0) {
                    const double t = -mfh_dot_exact(n - k, ck + k, 1, cj + k, 1) / ck[k];
                    for (int i = k; i < n; ++i) cj[i] = cj[i] + ck[i] * t;
                }
                E[j] = cj[k];
            }
            __syncthreads();
            if (k < n - 2) {
                double ek = norm_seq(n - kp1, E + kp1);
                if (ek != 0.0) {
                    if (E[kp1] < 0.0) ek = -ek;
                    for (int j = kp1 + lane; j < n; j += GW) {
                        double v = E[j] / ek;
                        if (j == kp1) v = v + 1.0;
                        E[j] = v;
                    }
                }
                __syncthreads();
                E[k] = -ek;
                if (ek != 0.0) {
                    // work[i] = sum_j a[i, j] e[j] (MulAdd in j order), then a[i, j] += work[i] (-e[j] / e[k+1])
                    for (int i = kp1 + lane; i < n; i += GW) {
                        double wk = 0.0;
                        for (int j = kp1; j < n; ++j) wk = wk + C[i + j * ldc] * E[j];
                        for (int j = kp1; j < n; ++j) C[i + j * ldc] = C[i + j * ldc] + wk * (-E[j] / E[kp1]);
                    }
                }
                __syncthreads();
            }
        }
        // The final bidiagonal matrix (:919-930).
        S[n - 1] = C[(n - 1) + (n - 1) * ldc];
        if (n >= 2) E[n - 2] = C[(n - 2) + (n - 1) * ldc];
        E[n - 1] = 0.0;
        // Generate U (:933-963).
        for (int i = lane; i < n; i += GW) U[i + (size_t)(n - 1) * ldu] = i == n - 1 ? 1.0 : 0.0;
        __syncthreads();
        for (int k = n - 2; k >= 0; --k) {
            double* uk = U + (size_t)k * ldu;
            if (S[k] != 0.0) {
                for (int j = k + 1 + lane; j < n; j += GW) {
                    double* uj = U + (size_t)j * ldu;
                    const double t = -mfh_dot_exact(n - k, uk + k, 1, uj + k, 1) / uk[k];
                    for (int i = k; i < n; ++i) uj[i] = uj[i] + uk[i] * t;
                }
                __syncthreads();
                for (int i = lane; i < n; i += GW) {
                    double v = 0.0;
                    if (i >= k) {
                        v = -uk[i];
                        if (i == k) v = v + 1.0;
                    }
                    uk[i] = v;
                }
            } else {
                for (int i = lane; i < n; i += GW) uk[i] = i == k ? 1.0 : 0.0;
            }
            __syncthreads();
        }

        // Singular values (:983-1232).  Every lane runs the same scalar
        // iteration on S and E; each lane rotates its own rows of U.
        int p = n;
        const int pp = n - 1;
        int iter = 0;
        long passes = 0;
        const long cap = svd_pass_cap(n);
        const double eps = 0x1p-52;
        bool ok = true;
        while (p > 0) {
            if (iter >= 1000 || ++passes > cap) {
                ok = false;
                break;
            }
            int k;
            for (k = p - 2; k >= 0; --k) {
                if (__builtin_fabs(E[k]) <= eps * (__builtin_fabs(S[k]) + __builtin_fabs(S[k + 1]))) {
                    E[k] = 0.0;
                    break;
                }
            }
            int kase;
            if (k == p - 2) {
                kase = 4;
            } else {
                int ks;
                for (ks = p - 1; ks > k; --ks) {
                    double t = 0.0;
                    if (ks != p - 1) t = t + __builtin_fabs(E[ks]);
                    if (ks != k + 1) t = t + __builtin_fabs(E[ks - 1]);
                    if (__builtin_fabs(S[ks]) <= eps * t) {
                        S[ks] = 0.0;
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
            if (kase == 1) {  // deflate a negligible s(p): rotations of V only, which is not formed
                f = E[p - 2];
                E[p - 2] = 0.0;
                for (int j = k; j < p - 1; ++j) {
                    const int l = p - 2 - j + k;
                    double t = S[l];
                    drotg(t, f, cs, sn);
                    S[l] = t;
                    if (l != k) {
                        f = -sn * E[l - 1];
                        E[l - 1] = E[l - 1] * cs;
                    }
                }
            } else if (kase == 2) {  // split at a negligible s(k)
                f = E[k - 1];
                E[k - 1] = 0.0;
                for (int j = k; j < p; ++j) {
                    double t = S[j];
                    drotg(t, f, cs, sn);
                    S[j] = t;
                    f = -sn * E[j];
                    E[j] = cs * E[j];
                    rot_rows(U, ldu, n, lane, j, k - 1, cs, sn);
                }
            } else if (kase == 3) {  // one QR step
                double scale = 0.0;
                scale = net_max(scale, __builtin_fabs(S[p - 1]));
                scale = net_max(scale, __builtin_fabs(S[p - 2]));
                scale = net_max(scale, __builtin_fabs(E[p - 2]));
                scale = net_max(scale, __builtin_fabs(S[k]));
                scale = net_max(scale, __builtin_fabs(E[k]));
                const double sp = S[p - 1] / scale;
                const double spm1 = S[p - 2] / scale;
                const double epm1 = E[p - 2] / scale;
                const double skk = S[k] / scale;
                const double ekk = E[k] / scale;
                const double bb = ((spm1 + sp) * (spm1 - sp) + epm1 * epm1) / 2.0;
                const double cc = (sp * epm1) * (sp * epm1);
                double shift = 0.0;
                if (bb != 0.0 || cc != 0.0) {
                    shift = __builtin_sqrt(bb * bb + cc);
                    if (bb < 0.0) shift = -shift;
                    shift = cc / (bb + shift);
                }
                f = (skk + sp) * (skk - sp) + shift;
                double g = skk * ekk;
                // The sweep's chain runs in registers: s(j) and e(j) are
                // carried from the step before (which last wrote them), and
                // s(j+1), e(j+1), not yet touched by this sweep, are read at
                // the top of the step, off the chain.  Every value and every
                // store is the reference's (:1134-1177), in its order.
                // The rotations of U's columns (j, j+1) carry column j+1 of
                // the lane's rows (it is column j of the next rotation), so a
                // step reads only column j+1, also at its top (:1167-1176).
                double sj = S[k], ej = E[k];
                double ux[2], uy[2];
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int i = lane + GW * r;
                    ux[r] = i < n ? U[i + (size_t)k * ldu] : 0.0;
                }
                for (int j = k; j < p - 1; ++j) {
                    const double sj1 = S[j + 1], ej1 = E[j + 1];
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        const int i = lane + GW * r;
                        uy[r] = i < n ? U[i + (size_t)(j + 1) * ldu] : 0.0;
                    }
                    drotg(f, g, cs, sn);
                    if (j != k) E[j - 1] = f;
                    f = cs * sj + sn * ej;
                    const double ejr = cs * ej - sn * sj;
                    g = sn * sj1;
                    const double sj1r = sj1 * cs;
                    drotg(f, g, cs, sn);
                    S[j] = f;
                    f = cs * ejr + sn * sj1r;
                    const double sj1s = -(sn * ejr) + cs * sj1r;
                    g = sn * ej1;
                    const double ej1r = ej1 * cs;
                    E[j] = ejr;
                    S[j + 1] = sj1s;
                    E[j + 1] = ej1r;
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        const int i = lane + GW * r;
                        if (i < n) U[i + (size_t)j * ldu] = cs * ux[r] + sn * uy[r];
                        ux[r] = cs * uy[r] - sn * ux[r];
                    }
                    sj = sj1s;
                    ej = ej1r;
                }
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int i = lane + GW * r;
                    if (i < n) U[i + (size_t)(p - 1) * ldu] = ux[r];
                }
                E[p - 2] = f;
                iter++;
            } else {  // convergence: make s(k) positive and move it into order
                if (S[k] < 0.0) S[k] = -S[k];
                while (k < pp) {
                    if (S[k] >= S[k + 1]) break;
                    const double t = S[k];
                    S[k] = S[k + 1];
                    S[k + 1] = t;
                    for (int i = lane; i < n; i += GW) {
                        const double x = U[i + (size_t)k * ldu];
                        U[i + (size_t)k * ldu] = U[i + (size_t)(k + 1) * ldu];
                        U[i + (size_t)(k + 1) * ldu] = x;
                    }
                    k++;
                }
                iter = 0;
                passes = 0;
                p--;
            }
        }
        __syncthreads();

        if (ok) {
            double* W = G.w[mi];
            for (int i = lane; i < n; i += GW) W[i] = S[i];
            // Eigenvectors: Lower/Trans solve of every column of U (:371-374).
            for (int c = lane; c < n; c += GW) {
                double* x = U + (size_t)c * ldu;
                for (int j = n - 1; j >= 0; --j) {
                    x[j] = x[j] / L[j + (size_t)j * ldb];
                    const double y = x[j];
                    for (int r = 0; r < j; ++r) x[r] = x[r] - L[j + (size_t)r * ldb] * y;
                }
            }
            __syncthreads();
        }
        if (ULDS) {  // a holds U: the eigenvectors, or the vectors reached when the iteration gave up
            for (int e = lane; e < n * n; e += GW) {
                const int j = e / n;
                const int i = e - j * n;
                A[i + (size_t)j * lda] = U[i + (size_t)j * ldu];
            }
        }
        G.info[mi] = ok ? 0 : n + 1;  // every lane stores the same value
    }
}

size_t lds_bytes(int max_n, int ldc, bool ulds)
{
    return ((size_t)((ulds ? 1 : 0) + (MFH_GEVD_C_IN_LDS ? 1 : 0)) * max_n * ldc + 2 * (size_t)max_n) * sizeof(double);
}

template <bool ULDS>
int launch_gevd(mfh_handle* h, int count, const GevdBatch& G, int max_n, int ldc)
{
    const size_t lds = lds_bytes(max_n, ldc, ULDS);
    MFH_HIP_CHECK(hipFuncSetAttribute((const void*)k_gevd_exact<ULDS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gevd_exact<ULDS>, GW, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > count) grid = count;
    int st = mfh_timing_begin(h, MFH_KERNEL_GEVD_EXACT, max_n);
    if (st) return st;
    hipLaunchKernelGGL(k_gevd_exact<ULDS>, dim3((unsigned)grid), dim3(GW), lds, h->stream, G, count, max_n, ldc);
    MFH_HIP_CHECK(hipGetLastError());
    return mfh_timing_end(h);
}

}  // namespace

// The handle's second buffer holds n_chol, info_chol and the order (count
// ints each), then the work counter, the order histogram and its cursors.
constexpr int kGevdHistOff = 64, kGevdCursorOff = 128, kGevdHeaderInts = 256;
static_assert(MfhOrder<MFH_SMALL_MAX_N>::keys <= kGevdCursorOff - kGevdHistOff &&
                  kGevdCursorOff + MfhOrder<MFH_SMALL_MAX_N>::keys <= kGevdHeaderInts,
              "order histogram");
static size_t gevd_ints(int count) { return ((size_t)count * sizeof(int) + 255) & ~(size_t)255; }
size_t mfh_gevd_aux_bytes(int count) { return 3 * gevd_ints(count) + kGevdHeaderInts * sizeof(int); }

int mfh_launch_gevd(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, double* const* b,
                    const int* ldb, double* const* w, int* info, int max_n)
{
    // C in HBM (when not in LDS): one max_n x ldc matrix per resident workgroup
    size_t cws_bytes = 0;
    if (!MFH_GEVD_C_IN_LDS) {
        const size_t lds = lds_bytes(max_n, max_n | 1, false);
        MFH_HIP_CHECK(hipFuncSetAttribute((const void*)k_gevd_exact<false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int per_cu = 0;
        MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gevd_exact<false>, GW, lds));
        if (per_cu < 1) per_cu = 1;
        cws_bytes = (size_t)per_cu * h->num_cus * max_n * (max_n | 1) * sizeof(double);
    }
    int st = mfh_aux_reserve(h, mfh_gevd_aux_bytes(count) + cws_bytes);
    if (st) return st;
    char* aux = (char*)h->aux;
    const size_t part = gevd_ints(count);
    int* header = (int*)(aux + 3 * part);
    GevdBatch G{n,      a,    lda,        b,    ldb, w, info, (int*)aux, (int*)(aux + part), (const int*)(aux + 2 * part),
                header, (double*)(aux + mfh_gevd_aux_bytes(count))};
    hipLaunchKernelGGL(k_gevd_check, dim3((unsigned)((count + GW - 1) / GW)), dim3(GW), 0, h->stream, G, count,
                       max_n);
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(256), 0, h->stream, header, kGevdHeaderInts);
    mfh_launch_order<MFH_SMALL_MAX_N>(h->stream, n, count, header + kGevdHistOff, header + kGevdCursorOff,
                                      (int*)(aux + 2 * part));
    MFH_HIP_CHECK(hipGetLastError());
    // B = L L^T with the MatFlat-order Cholesky whatever the handle's math
    // mode; matrices with invalid arguments have order 0 there and are skipped.
    const int mode = h->math_mode;
    h->math_mode = MFH_MATH_EXACT;
    MfhCholBatch B{G.n_chol, b, ldb, nullptr, G.info_chol};
    st = mfh_launch_potrf(h, count, B, max_n, false);
    h->math_mode = mode;
    if (st) return st;
    const int ldc = max_n | 1;  // odd: lanes walking rows of C hit distinct LDS banks
    // U is built in a, as the reference builds it, and only C is in LDS:
    // 33.8 KB per problem at max_n = 64, four problems per CU instead of two
    // with U in LDS too (10,000 pairs, n ~ U{8..64}: 19.6 against 29.7 ms;
    // the sweep's rotations read one column of U per step, off the chain).
    // MFH_GEVD_U_IN_LDS=1 builds the LDS variant for comparison.
#ifndef MFH_GEVD_U_IN_LDS
#define MFH_GEVD_U_IN_LDS 0
#endif
    if (MFH_GEVD_U_IN_LDS && lds_bytes(max_n, ldc, true) <= 160 * 1024) return launch_gevd<true>(h, count, G, max_n, ldc);
    return launch_gevd<false>(h, count, G, max_n, ldc);
}
