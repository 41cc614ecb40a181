// mfh_gemm.hip -- variable-size batched GEMM (FP64 and complex128) for
// gfx950 in MatFlat's exact summation order.
//
// Blas.MulMatMat (MatFlat/Blas.MulMatMat.cs:226-368 double, :407-648
// Complex) computes every C element as one dot product of length k:
// Internals.Dot's pairwise order for double (Internals.cs:265-292) and a
// sequential sum of BCL Complex products for complex (Dot_NN/CN/NC/CC,
// :598-648).  A dot product's order is a per-element property, so the
// kernel keeps it while organising the work like a GEMM:
//  * a workgroup of 64 threads owns a 32x32 tile of C; each thread owns a
//    4x4 register block (rows r+8i, columns c+8j) and its 16 running sums;
//  * op(A) and op(B) are staged through LDS in chunks of 16 k-values
//    (k-major), so each operand byte is read from L2 once per tile and each
//    LDS value feeds 4 products;
//  * the tiles of all products form one work list (device-side prefix sum
//    of per-product tile counts) that persistent workgroups drain through an
//    atomic counter, so one launch covers any mix of shapes.
// No product is ever contracted into an FMA (-ffp-contract=off), so with
// alpha = 1 and beta = 0 the output is bit-identical to MatFlat's.
#include "mfh_device.h"
#include "mfh_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int TM = 32;
constexpr int TN = 32;
constexpr int KC = 16;
constexpr int GT = 64;  // threads per workgroup

struct GemmWs {
    int* tiles;    // per-product tile count, then exclusive scan (count + 1)
    int* counter;  // [0] work counter
};

__device__ inline bool is_trans(int t) { return t == MFH_TRANS || t == MFH_CONJTRANS; }
__device__ inline bool is_conj(int t) { return t == MFH_CONJTRANS || t == MFH_CONJNOTRANS; }

__global__ void k_gemm_validate(MfhGemmBatch B, int count, GemmWs ws, int max_m, int max_n, int cplx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) ws.counter[0] = 0;
    if (i >= count) return;
    const int ta = B.transa[i], tb = B.transb[i];
    const int m = B.m[i], n = B.n[i], k = B.k[i];
    int bad = 0;
    const bool ta_ok = cplx ? (ta >= MFH_NOTRANS && ta <= MFH_CONJNOTRANS) : (ta == MFH_NOTRANS || ta == MFH_TRANS);
    const bool tb_ok = cplx ? (tb >= MFH_NOTRANS && tb <= MFH_CONJNOTRANS) : (tb == MFH_NOTRANS || tb == MFH_TRANS);
    if (!ta_ok) bad = -3;
    else if (!tb_ok) bad = -4;
    else if (m < 1 || m > max_m) bad = -5;
    else if (n < 1 || n > max_n) bad = -6;
    else if (k < 1) bad = -7;
    else if (B.a[i] == nullptr) bad = -9;
    else if (B.lda[i] < (is_trans(ta) ? k : m)) bad = -10;
    else if (B.b[i] == nullptr) bad = -11;
    else if (B.ldb[i] < (is_trans(tb) ? n : k)) bad = -12;
    else if (B.c[i] == nullptr) bad = -14;
    else if (B.ldc[i] < m) bad = -15;
    else if ((const double*)B.c[i] == B.a[i] || (const double*)B.c[i] == B.b[i]) bad = -1;
    B.info[i] = bad;
    ws.tiles[i] = bad ? 0 : ((m + TM - 1) / TM) * ((n + TN - 1) / TN);
}

// Exclusive scan of count ints in place (one workgroup of 1024 threads),
// writing the total to tiles[count].
__global__ __launch_bounds__(1024) void k_scan(int* v, int count)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (count + 1023) / 1024;
    const int lo = t * per;
    const int hi = lo + per < count ? lo + per : count;
    int s = 0;
    for (int i = lo; i < hi; ++i) s += v[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int add = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int run = (t == 0) ? 0 : part[t - 1];
    for (int i = lo; i < hi; ++i) {
        int x = v[i];
        v[i] = run;
        run += x;
    }
    if (t == 1023) v[count] = part[1023];
}

__device__ inline int find_product(const int* scan, int count, int tile)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (scan[mid] <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Loads a KC x 32 chunk of op(X) into LDS as S[kk][r] (k-major).  op(X) is
// rows x k; not transposed: element (r, kk) at X[r + kk*ld], transposed:
// X[kk + r*ld].  Elements outside the matrix are stored as zero and never
// contribute to a stored C element.
template <typename V>
__device__ inline void stage(V (*S)[TM + 1], const V* X, int ld, bool trans, int r0, int rows, int k0, int kend)
{
    for (int e = threadIdx.x; e < KC * TM; e += GT) {
        int r, kk;
        if (trans) {
            kk = e % KC;
            r = e / KC;
        } else {
            r = e % TM;
            kk = e / TM;
        }
        const int gr = r0 + r, gk = k0 + kk;
        V v{};
        if (gr < rows && gk < kend) v = trans ? X[gk + (size_t)gr * ld] : X[gr + (size_t)gk * ld];
        S[kk][r] = v;
    }
}

__global__ __launch_bounds__(GT) void k_dgemm_exact(MfhGemmBatch B, int count, GemmWs ws)
{
    __shared__ double As[KC][TM + 1];
    __shared__ double Bs[KC][TN + 1];
    __shared__ int s_tile;
    const int t = threadIdx.x;
    const int tr = t % 8, tc = t / 8;
    const int total = ws.tiles[count];
    for (;;) {
        if (t == 0) s_tile = atomicAdd(ws.counter, 1);
        __syncthreads();
        const int tile = s_tile;
        __syncthreads();
        if (tile >= total) break;
        const int p = find_product(ws.tiles, count, tile);
        const int local = tile - ws.tiles[p];
        const int m = B.m[p], n = B.n[p], k = B.k[p];
        const int mt = (m + TM - 1) / TM;
        const int i0 = (local % mt) * TM, j0 = (local / mt) * TN;
        const bool ta = B.transa[p] == MFH_TRANS, tb = B.transb[p] == MFH_TRANS;
        const double* A = B.a[p];
        const double* Bm = B.b[p];
        const int lda = B.lda[p], ldb = B.ldb[p];

        double s[4][4], pend[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                s[a][c] = 0.0;
                pend[a][c] = 0.0;
            }
        // op(B) is k x n; staged as Bs[kk][col] = op(B)(kk, col): for
        // NoTrans that is B[kk + col*ldb] (the "transposed" load pattern of
        // a rows x k operand), for Trans B[col + kk*ldb].
        const int k0 = k & 1;  // odd k: element 0 first (Internals.cs:271-276)
        if (k0) {
            const int a_r0 = 0;
            (void)a_r0;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int gi = i0 + tr + 8 * a;
                double av = 0.0;
                if (gi < m) av = ta ? A[(size_t)gi * lda] : A[gi];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int gj = j0 + tc + 8 * c;
                    double bv = 0.0;
                    if (gj < n) bv = tb ? Bm[gj] : Bm[(size_t)gj * ldb];
                    s[a][c] = av * bv;
                }
            }
        }
        for (int kc = k0; kc < k; kc += KC) {
            __syncthreads();
            stage<double>(As, A, lda, ta, i0, m, kc, k);
            stage<double>(Bs, Bm, ldb, !tb, j0, n, kc, k);
            __syncthreads();
            const int len = (k - kc) < KC ? (k - kc) : KC;  // even
            for (int q = 0; q < len; q += 2) {
                double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    a0[a] = As[q][tr + 8 * a];
                    a1[a] = As[q + 1][tr + 8 * a];
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    b0[c] = Bs[q][tc + 8 * c];
                    b1[c] = Bs[q + 1][tc + 8 * c];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        double p0 = a0[a] * b0[c];
                        double p1 = a1[a] * b1[c];
                        s[a][c] = s[a][c] + (p0 + p1);
                    }
            }
        }
        (void)pend;
        const double alpha = B.alpha[p], beta = B.beta[p];
        double* C = B.c[p];
        const int ldc = B.ldc[p];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gj = j0 + tc + 8 * c;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int gi = i0 + tr + 8 * a;
                if (gi < m && gj < n) {
                    double v = (alpha == 1.0) ? s[a][c] : alpha * s[a][c];
                    double* cp = C + gi + (size_t)gj * ldc;
                    if (beta != 0.0) v = v + beta * *cp;
                    *cp = v;
                }
            }
        }
    }
}

struct Cplx {
    double re, im;
};

// BCL Complex multiply: re = a*c - b*d, im = b*c + a*d.
__device__ inline Cplx cmul(double a, double b, double c, double d)
{
    return Cplx{a * c - b * d, b * c + a * d};
}

// Products of Dot_NN / Dot_CN / Dot_NC / Dot_CC (Blas.MulMatMat.cs:598-648,
// Internals.MulConj/MulConjConj, Internals.cs:21-39) with x = op(A)(i,kk),
// y = op(B)(kk,j).
__device__ inline Cplx zprod(int kind, Cplx x, Cplx y)
{
    switch (kind) {
    case 0: return cmul(x.re, x.im, y.re, y.im);
    case 1: {  // MulConj(x, y)
        double a = x.re, b = -x.im, c = y.re, d = y.im;
        return Cplx{a * c - b * d, a * d + b * c};
    }
    case 2: {  // MulConj(y, x)
        double a = y.re, b = -y.im, c = x.re, d = x.im;
        return Cplx{a * c - b * d, a * d + b * c};
    }
    default: {  // MulConjConj(x, y)
        double a = x.re, b = -x.im, c = y.re, d = -y.im;
        return Cplx{a * c - b * d, a * d + b * c};
    }
    }
}

template <int KIND>
__device__ inline void zgemm_tile(Cplx (*As)[TM + 1], Cplx (*Bs)[TN + 1], Cplx (&s)[4][4], int len, int tr, int tc)
{
    for (int q = 0; q < len; ++q) {
        Cplx av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = As[q][tr + 8 * a];
#pragma unroll
        for (int c = 0; c < 4; ++c) bv[c] = Bs[q][tc + 8 * c];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                Cplx pr = zprod(KIND, av[a], bv[c]);
                s[a][c].re = s[a][c].re + pr.re;
                s[a][c].im = s[a][c].im + pr.im;
            }
    }
}

__global__ __launch_bounds__(GT) void k_zgemm_exact(MfhGemmBatch B, int count, GemmWs ws)
{
    __shared__ Cplx As[KC][TM + 1];
    __shared__ Cplx Bs[KC][TN + 1];
    __shared__ int s_tile;
    const int t = threadIdx.x;
    const int tr = t % 8, tc = t / 8;
    const int total = ws.tiles[count];
    for (;;) {
        if (t == 0) s_tile = atomicAdd(ws.counter, 1);
        __syncthreads();
        const int tile = s_tile;
        __syncthreads();
        if (tile >= total) break;
        const int p = find_product(ws.tiles, count, tile);
        const int local = tile - ws.tiles[p];
        const int m = B.m[p], n = B.n[p], k = B.k[p];
        const int mt = (m + TM - 1) / TM;
        const int i0 = (local % mt) * TM, j0 = (local / mt) * TN;
        const int opa = B.transa[p], opb = B.transb[p];
        const bool ta = is_trans(opa), tb = is_trans(opb);
        const int kind = is_conj(opa) ? (is_conj(opb) ? 3 : 1) : (is_conj(opb) ? 2 : 0);
        const Cplx* A = (const Cplx*)B.a[p];
        const Cplx* Bm = (const Cplx*)B.b[p];
        const int lda = B.lda[p], ldb = B.ldb[p];
        Cplx s[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) s[a][c] = Cplx{0.0, 0.0};
        for (int kc = 0; kc < k; kc += KC) {
            __syncthreads();
            stage<Cplx>(As, A, lda, ta, i0, m, kc, k);
            stage<Cplx>(Bs, Bm, ldb, !tb, j0, n, kc, k);
            __syncthreads();
            const int len = (k - kc) < KC ? (k - kc) : KC;
            switch (kind) {
            case 0: zgemm_tile<0>(As, Bs, s, len, tr, tc); break;
            case 1: zgemm_tile<1>(As, Bs, s, len, tr, tc); break;
            case 2: zgemm_tile<2>(As, Bs, s, len, tr, tc); break;
            default: zgemm_tile<3>(As, Bs, s, len, tr, tc); break;
            }
        }
        const double ar = B.alpha[2 * p], ai = B.alpha[2 * p + 1];
        const double br = B.beta[2 * p], bi = B.beta[2 * p + 1];
        const bool alpha_one = ar == 1.0 && ai == 0.0;
        const bool beta_zero = br == 0.0 && bi == 0.0;
        Cplx* C = (Cplx*)B.c[p];
        const int ldc = B.ldc[p];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gj = j0 + tc + 8 * c;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int gi = i0 + tr + 8 * a;
                if (gi < m && gj < n) {
                    Cplx v = s[a][c];
                    if (!alpha_one) v = cmul(ar, ai, v.re, v.im);
                    Cplx* cp = C + gi + (size_t)gj * ldc;
                    if (!beta_zero) {
                        Cplx u = cmul(br, bi, cp->re, cp->im);
                        v.re = v.re + u.re;
                        v.im = v.im + u.im;
                    }
                    *cp = v;
                }
            }
        }
    }
}

}  // namespace

int mfh_launch_gemm(mfh_handle* h, int count, const MfhGemmBatch& B, int max_m, int max_n, int max_k, bool cplx)
{
    (void)max_k;
    size_t bytes = sizeof(int) * ((size_t)count + 2 + 64);
    int st = mfh_ws_reserve(h, bytes);
    if (st) return st;
    GemmWs ws{(int*)h->ws, (int*)h->ws + count + 1};
    hipLaunchKernelGGL(k_gemm_validate, dim3((count + 255) / 256), dim3(256), 0, h->stream, B, count, ws, max_m,
                       max_n, cplx ? 1 : 0);
    MFH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, h->stream, ws.tiles, count);
    MFH_HIP_CHECK(hipGetLastError());
    int per_cu = 0;
    if (cplx) MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_zgemm_exact, GT, 0));
    else MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dgemm_exact, GT, 0));
    if (per_cu < 1) per_cu = 1;
    long maxtiles = (long)count * ((max_m + TM - 1) / TM) * ((max_n + TN - 1) / TN);
    long grid = (long)per_cu * h->num_cus;
    if (grid > maxtiles) grid = maxtiles;
    if (grid < 1) grid = 1;
    if (cplx) hipLaunchKernelGGL(k_zgemm_exact, dim3((unsigned)grid), dim3(GT), 0, h->stream, B, count, ws);
    else hipLaunchKernelGGL(k_dgemm_exact, dim3((unsigned)grid), dim3(GT), 0, h->stream, B, count, ws);
    MFH_HIP_CHECK(hipGetLastError());
    return MFH_OK;
}
