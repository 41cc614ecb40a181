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
//    of per-product tile counts, then a tile -> product map) that
//    persistent workgroups walk with a fixed stride, so one launch covers
//    any mix of shapes with no atomics and no search per tile.
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
    int* tiles;  // exclusive prefix of per-product tile counts (count + 1)
    int* cnt;    // per-product tile count (count)
    int* bsum;   // tile total per chunk of SCAN_CHUNK products
    int* map;    // tile -> product (tiles[count] entries)
};

// The tile list is built by two kernels over chunks of SCAN_CHUNK products:
// validate + chunk-local scan, then chunk offsets + tile -> product map.
constexpr int SCAN_T = 256;
constexpr int SCAN_E = 16;
constexpr int SCAN_CHUNK = SCAN_T * SCAN_E;

// XCD-aware walk of the tile list.  Workgroups are dealt round-robin to the
// eight XCDs of the device, each with its own L2: workgroup w runs on XCD
// w % 8.  Walking the list with one global stride puts the tiles of a
// product -- which read the same rows of op(A) and columns of op(B) -- on up
// to four XCDs, and each of those L2s fetches the operands from HBM again.
// Here the list is cut into eight contiguous ranges and the workgroups of
// XCD x walk range x with their own stride, so a product's tiles share one
// L2 (only the products that straddle a range boundary are split).
constexpr int NXCD = 8;

struct TileWalk {
    int next, end, stride;
};

__device__ inline TileWalk tile_walk(int total, int per_group, int sub)
{
    const int groups = (int)gridDim.x;
    const int nx = (groups % NXCD == 0) ? NXCD : 1;  // the host makes the grid a multiple of 8
    const int x = (int)blockIdx.x % nx, r = (int)blockIdx.x / nx;
    const int first = (int)((long)total * x / nx), end = (int)((long)total * (x + 1) / nx);
    return TileWalk{first + r * per_group + sub, end, (groups / nx) * per_group};
}

// Grid size for the XCD-aware walk: a multiple of the XCD count once there
// are at least that many workgroups.
inline long xcd_grid(long grid) { return grid >= NXCD ? grid / NXCD * NXCD : grid; }

__device__ inline bool is_trans(int t) { return t == MFH_TRANS || t == MFH_CONJTRANS; }
__device__ inline bool is_conj(int t) { return t == MFH_CONJTRANS || t == MFH_CONJNOTRANS; }

// Checks product i in Blas.MulMatMat's argument order (Blas.MulMatMat.cs:
// 230-304), stores info[i] and returns its number of ts x ts tiles of C (0
// when it is invalid).
__device__ inline int gemm_validate(const MfhGemmBatch& B, int i, int max_m, int max_n, int cplx, int ts)
{
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
    return bad ? 0 : ((m + ts - 1) / ts) * ((n + ts - 1) / ts);
}

// Exclusive scan of one int per thread over a workgroup of SCAN_T threads.
__device__ inline int block_exclusive_scan(int v, int* part, int& total)
{
    const int t = threadIdx.x;
    part[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        const int add = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    total = part[SCAN_T - 1];
    const int excl = part[t] - v;
    __syncthreads();
    return excl;
}

__global__ __launch_bounds__(SCAN_T) void k_gemm_validate_scan(MfhGemmBatch B, int count, GemmWs ws, int max_m,
                                                               int max_n, int cplx, int ts)
{
    __shared__ int part[SCAN_T];
    const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_E;
    int c[SCAN_E];
    int sum = 0;
#pragma unroll
    for (int e = 0; e < SCAN_E; ++e) {
        c[e] = base + e < count ? gemm_validate(B, base + e, max_m, max_n, cplx, ts) : 0;
        sum += c[e];
    }
    int total;
    int run = block_exclusive_scan(sum, part, total);
#pragma unroll
    for (int e = 0; e < SCAN_E; ++e)
        if (base + e < count) {
            ws.tiles[base + e] = run;
            ws.cnt[base + e] = c[e];
            run += c[e];
        }
    if (threadIdx.x == 0) ws.bsum[blockIdx.x] = total;
}

// Adds the chunk offset (sum of the preceding chunks' totals) to each
// product's prefix, writes the grand total to tiles[count] and fills
// map[t] = p for every tile t of product p.
__global__ __launch_bounds__(SCAN_T) void k_gemm_map(GemmWs ws, int count, int nchunks)
{
    __shared__ int part[SCAN_T];
    int before = 0, all = 0;
    for (int c = threadIdx.x; c < nchunks; c += SCAN_T) {
        const int v = ws.bsum[c];
        all += v;
        if (c < (int)blockIdx.x) before += v;
    }
    int sum_before, sum_all;
    block_exclusive_scan(before, part, sum_before);
    block_exclusive_scan(all, part, sum_all);
    const int base = blockIdx.x * SCAN_CHUNK;
    if (blockIdx.x == 0 && threadIdx.x == 0) ws.tiles[count] = sum_all;
    for (int i = threadIdx.x; i < SCAN_CHUNK; i += SCAN_T) {
        const int p = base + i;
        if (p >= count) break;
        const int t0 = ws.tiles[p] + sum_before;
        ws.tiles[p] = t0;
        for (int t = 0; t < ws.cnt[p]; ++t) ws.map[t0 + t] = p;
    }
}

// Loads a KC x 32 chunk of op(X) into LDS as S[kk][r] (k-major).  op(X) is
// rows x k; not transposed: element (r, kk) at X[r + kk*ld], transposed:
// X[kk + r*ld].  Elements outside the matrix are stored as zero and never
// contribute to a stored C element.
template <typename V>
__device__ inline void stage(V (*S)[TM + 1], const V* X, int ld, bool trans, int r0, int rows, int k0, int kend)
{
    // every load of the chunk goes out before the first LDS store: a loop
    // of load, wait and store per element serialised PER round trips
    constexpr int PER = KC * TM / GT;
    static_assert(KC * TM % GT == 0, "chunk must divide over the workgroup");
    V v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = threadIdx.x + u * GT;
        const int r = trans ? e / KC : e % TM;
        const int kk = trans ? e % KC : e / TM;
        const int gr = r0 + r, gk = k0 + kk;
        v[u] = V{};
        if (gr < rows && gk < kend) v[u] = trans ? X[gk + (size_t)gr * ld] : X[gr + (size_t)gk * ld];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = threadIdx.x + u * GT;
        const int r = trans ? e / KC : e % TM;
        const int kk = trans ? e % KC : e / TM;
        S[kk][r] = v[u];
    }
}

// Held to four wavefronts per SIMD (128 registers, 80 bytes spilled) instead
// of the three its natural 148 VGPRs allow: a tile's chunk loads and its dot
// chains are latency-bound, and the extra wavefront hides more of them than
// the spill costs (200,000 products: 19.45 -> 15.05 ms).
__global__ __launch_bounds__(GT, 4) void k_dgemm_exact(MfhGemmBatch B, int count, GemmWs ws)
{
    __shared__ double As[KC][TM + 1];
    __shared__ double Bs[KC][TN + 1];
    const int t = threadIdx.x;
    const int tr = t % 8, tc = t / 8;
    const TileWalk walk = tile_walk(ws.tiles[count], 1, 0);
    for (int tile = walk.next; tile < walk.end; tile += walk.stride) {
        const int p = ws.map[tile];
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

// Held to two wavefronts per SIMD (256 registers, 72 bytes spilled): left
// alone the compiler took 256 VGPRs and AGPRs beyond them, one wavefront per
// SIMD (200,000 products: 64.6 -> 43.0 ms).  Three (168 registers) spills
// 420 bytes per lane.
__global__ __launch_bounds__(GT, 2) void k_zgemm_exact(MfhGemmBatch B, int count, GemmWs ws)
{
    __shared__ Cplx As[KC][TM + 1];
    __shared__ Cplx Bs[KC][TN + 1];
    const int t = threadIdx.x;
    const int tr = t % 8, tc = t / 8;
    const TileWalk walk = tile_walk(ws.tiles[count], 1, 0);
    for (int tile = walk.next; tile < walk.end; tile += walk.stride) {
        const int p = ws.map[tile];
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

// ---- FAST mode: FP64 matrix cores ------------------------------------------
// With MFH_MATH_FAST the products run on v_mfma_f64_16x16x4f64 instead of
// MatFlat's dot order; results agree with MatFlat to rounding (|error| <=
// k eps |op(A)| |op(B)| per element).  One wavefront owns a 32x32 tile of C
// and computes its transpose D = C^T = op(B)^T op(A)^T as 2x2 (or fewer, at
// the matrix edge) 16x16 blocks, so that each lane's result column is a run
// of 16 consecutive rows of C and the stores coalesce.  Operands go straight
// from global memory into the MFMA operand registers -- lane l holds
// element (row l%16, k l/16) of its 16x4 fragment -- with no LDS and no
// barriers: the four wavefronts of a workgroup are independent workers on
// the strided tile list.  Rows outside the matrix load a clamped address
// and contribute zero; k runs in steps of 4 with a masked tail step.
typedef double mfh_d4 __attribute__((ext_vector_type(4)));

constexpr int FW = 4;  // wavefronts per workgroup
#ifndef MFH_GEMM_U
#define MFH_GEMM_U 2
#endif

// One operand side: NB fragments of 16 rows each.  Element (r, kk) of op(X)
// is X[kk + r*ld] when kmajor, else X[r + kk*ld].
template <typename V, int NB>
struct Side {
    const V* p[NB];  // this lane's element at the current k step
    bool ok[NB];     // lane row inside the matrix
    long inc;        // elements per k step of 4
};

template <typename V, int NB>
__device__ __forceinline__ Side<V, NB> side_at(const V* X, int ld, bool kmajor, int r0, int rows, int lane)
{
    Side<V, NB> S;
    const int kl = lane >> 4;
#pragma unroll
    for (int f = 0; f < NB; ++f) {
        int r = r0 + 16 * f + (lane & 15);
        S.ok[f] = r < rows;
        r = r < rows ? r : rows - 1;
        S.p[f] = X + (kmajor ? ((long)kl + (long)r * ld) : ((long)r + (long)kl * ld));
    }
    S.inc = kmajor ? 4 : 4L * ld;
    return S;
}

__device__ __forceinline__ double masked(double v, bool ok) { return ok ? v : 0.0; }
__device__ __forceinline__ Cplx masked(Cplx v, bool ok, double sgn)
{
    return Cplx{ok ? v.re : 0.0, ok ? sgn * v.im : 0.0};
}

// acc[bi][bj] += op(B)^T-fragment(bj) x op(A)^T-fragment(bi) over `steps`
// k steps of 4, in groups of U steps.  The loads of group g+1 are issued
// before the MFMAs of group g (the last group re-loads itself rather than
// run past the matrix).
template <int NBI, int NBJ, int U>
struct DGroup {
    double a[U][NBI], b[U][NBJ];
};

template <int NBI, int NBJ, int U>
__device__ __forceinline__ void dload(DGroup<NBI, NBJ, U>& G, const Side<double, NBI>& SA,
                                      const Side<double, NBJ>& SB, long step)
{
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int f = 0; f < NBI; ++f) G.a[u][f] = masked(SA.p[f][(step + u) * SA.inc], SA.ok[f]);
#pragma unroll
        for (int f = 0; f < NBJ; ++f) G.b[u][f] = masked(SB.p[f][(step + u) * SB.inc], SB.ok[f]);
    }
}

template <int NBI, int NBJ, int U>
__device__ __forceinline__ void dmfma(const DGroup<NBI, NBJ, U>& G, mfh_d4 (&acc)[NBI][NBJ])
{
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < NBI; ++i)
#pragma unroll
            for (int j = 0; j < NBJ; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(G.b[u][j], G.a[u][i], acc[i][j], 0, 0, 0);
}

template <int NBI, int NBJ, int U>
__device__ __forceinline__ void dmain(Side<double, NBI>& SA, Side<double, NBJ>& SB, mfh_d4 (&acc)[NBI][NBJ], int steps)
{
    const int groups = steps / U;
    if (groups > 0) {
        DGroup<NBI, NBJ, U> cur, nxt;
        dload(cur, SA, SB, 0);
        for (int g = 0; g < groups; ++g) {
            const int ng = g + 1 < groups ? g + 1 : g;
            dload(nxt, SA, SB, (long)ng * U);
            dmfma(cur, acc);
            cur = nxt;
        }
    }
    DGroup<NBI, NBJ, 1> one;
    for (int s = groups * U; s < steps; ++s) {
        dload(one, SA, SB, s);
        dmfma(one, acc);
    }
#pragma unroll
    for (int f = 0; f < NBI; ++f) SA.p[f] += (long)steps * SA.inc;
#pragma unroll
    for (int f = 0; f < NBJ; ++f) SB.p[f] += (long)steps * SB.inc;
}

template <int NBI, int NBJ>
__device__ __forceinline__ void dtile(const MfhGemmBatch& B, int p, int i0, int j0, int lane)
{
    const int m = B.m[p], n = B.n[p], k = B.k[p];
    const bool ta = B.transa[p] == MFH_TRANS, tb = B.transb[p] == MFH_TRANS;
    const int lda = B.lda[p], ldb = B.ldb[p];
    // op(A)(i, kk): NoTrans A[i + kk*lda], Trans A[kk + i*lda];
    // op(B)(kk, j): NoTrans B[kk + j*ldb], Trans B[j + kk*ldb].
    Side<double, NBI> SA = side_at<double, NBI>(B.a[p], lda, ta, i0, m, lane);
    Side<double, NBJ> SB = side_at<double, NBJ>(B.b[p], ldb, !tb, j0, n, lane);
    mfh_d4 acc[NBI][NBJ];
#pragma unroll
    for (int i = 0; i < NBI; ++i)
#pragma unroll
        for (int j = 0; j < NBJ; ++j) acc[i][j] = mfh_d4{0.0, 0.0, 0.0, 0.0};
    const int full = k >> 2, tail = k & 3;
    dmain<NBI, NBJ, MFH_GEMM_U>(SA, SB, acc, full);
    if (tail) {
        const bool kin = (lane >> 4) < tail;
        double av[NBI], bv[NBJ];
#pragma unroll
        for (int f = 0; f < NBI; ++f) av[f] = masked(kin ? SA.p[f][0] : 0.0, SA.ok[f]);
#pragma unroll
        for (int f = 0; f < NBJ; ++f) bv[f] = masked(kin ? SB.p[f][0] : 0.0, SB.ok[f]);
#pragma unroll
        for (int i = 0; i < NBI; ++i)
#pragma unroll
            for (int j = 0; j < NBJ; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
    }
    const double alpha = B.alpha[p], beta = B.beta[p];
    double* C = B.c[p];
    const long ldc = B.ldc[p];
#pragma unroll
    for (int i = 0; i < NBI; ++i) {
        const int gi = i0 + 16 * i + (lane & 15);
#pragma unroll
        for (int j = 0; j < NBJ; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gj = j0 + 16 * j + (lane >> 4) + 4 * r;
                if (gi < m && gj < n) {
                    double* cp = C + gi + gj * ldc;
                    double v = alpha * acc[i][j][r];
                    if (beta != 0.0) v = __builtin_fma(beta, *cp, v);
                    *cp = v;
                }
            }
        }
    }
}

// Complex: re += Bre Are - Bim Aim, im += Bre Aim + Bim Are, four real MFMAs
// per block; a conjugated operand has its imaginary part negated on load.
template <int NBI, int NBJ>
__device__ __forceinline__ void zstep(const Cplx (&av)[NBI], const Cplx (&bv)[NBJ], mfh_d4 (&re)[NBI][NBJ],
                                      mfh_d4 (&im)[NBI][NBJ])
{
#pragma unroll
    for (int i = 0; i < NBI; ++i)
#pragma unroll
        for (int j = 0; j < NBJ; ++j) {
            re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].re, av[i].re, re[i][j], 0, 0, 0);
            im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].re, av[i].im, im[i][j], 0, 0, 0);
            re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bv[j].im, av[i].im, re[i][j], 0, 0, 0);
            im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].im, av[i].re, im[i][j], 0, 0, 0);
        }
}

template <int NBI, int NBJ>
__device__ __forceinline__ void ztile(const MfhGemmBatch& B, int p, int i0, int j0, int lane)
{
    const int m = B.m[p], n = B.n[p], k = B.k[p];
    const int opa = B.transa[p], opb = B.transb[p];
    const int lda = B.lda[p], ldb = B.ldb[p];
    const double sa = is_conj(opa) ? -1.0 : 1.0, sb = is_conj(opb) ? -1.0 : 1.0;
    Side<Cplx, NBI> SA = side_at<Cplx, NBI>((const Cplx*)B.a[p], lda, is_trans(opa), i0, m, lane);
    Side<Cplx, NBJ> SB = side_at<Cplx, NBJ>((const Cplx*)B.b[p], ldb, !is_trans(opb), j0, n, lane);
    mfh_d4 re[NBI][NBJ], im[NBI][NBJ];
#pragma unroll
    for (int i = 0; i < NBI; ++i)
#pragma unroll
        for (int j = 0; j < NBJ; ++j) {
            re[i][j] = mfh_d4{0.0, 0.0, 0.0, 0.0};
            im[i][j] = mfh_d4{0.0, 0.0, 0.0, 0.0};
        }
    const int full = k >> 2, tail = k & 3;
    constexpr int U = 2;
    int s = 0;
    for (; s + U <= full; s += U) {
        Cplx av[U][NBI], bv[U][NBJ];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int f = 0; f < NBI; ++f) av[u][f] = masked(SA.p[f][u * SA.inc], SA.ok[f], sa);
#pragma unroll
            for (int f = 0; f < NBJ; ++f) bv[u][f] = masked(SB.p[f][u * SB.inc], SB.ok[f], sb);
        }
#pragma unroll
        for (int f = 0; f < NBI; ++f) SA.p[f] += U * SA.inc;
#pragma unroll
        for (int f = 0; f < NBJ; ++f) SB.p[f] += U * SB.inc;
#pragma unroll
        for (int u = 0; u < U; ++u) zstep<NBI, NBJ>(av[u], bv[u], re, im);
    }
    for (; s < full; ++s) {
        Cplx av[NBI], bv[NBJ];
#pragma unroll
        for (int f = 0; f < NBI; ++f) av[f] = masked(SA.p[f][0], SA.ok[f], sa);
#pragma unroll
        for (int f = 0; f < NBJ; ++f) bv[f] = masked(SB.p[f][0], SB.ok[f], sb);
#pragma unroll
        for (int f = 0; f < NBI; ++f) SA.p[f] += SA.inc;
#pragma unroll
        for (int f = 0; f < NBJ; ++f) SB.p[f] += SB.inc;
        zstep<NBI, NBJ>(av, bv, re, im);
    }
    if (tail) {
        const bool kin = (lane >> 4) < tail;
        const Cplx zero{0.0, 0.0};
        Cplx av[NBI], bv[NBJ];
#pragma unroll
        for (int f = 0; f < NBI; ++f) av[f] = masked(kin ? SA.p[f][0] : zero, SA.ok[f], sa);
#pragma unroll
        for (int f = 0; f < NBJ; ++f) bv[f] = masked(kin ? SB.p[f][0] : zero, SB.ok[f], sb);
        zstep<NBI, NBJ>(av, bv, re, im);
    }
    const double ar = B.alpha[2 * p], ai = B.alpha[2 * p + 1];
    const double br = B.beta[2 * p], bi = B.beta[2 * p + 1];
    const bool beta_zero = br == 0.0 && bi == 0.0;
    Cplx* C = (Cplx*)B.c[p];
    const long ldc = B.ldc[p];
#pragma unroll
    for (int i = 0; i < NBI; ++i) {
        const int gi = i0 + 16 * i + (lane & 15);
#pragma unroll
        for (int j = 0; j < NBJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gj = j0 + 16 * j + (lane >> 4) + 4 * r;
                if (gi < m && gj < n) {
                    Cplx* cp = C + gi + gj * ldc;
                    const double xr = re[i][j][r], xi = im[i][j][r];
                    double vr = ar * xr - ai * xi, vi = ar * xi + ai * xr;
                    if (!beta_zero) {
                        const Cplx c0 = *cp;
                        vr += br * c0.re - bi * c0.im;
                        vi += br * c0.im + bi * c0.re;
                    }
                    *cp = Cplx{vr, vi};
                }
            }
    }
}

// Tile edge of the FAST kernels: 32 (2x2 MFMA blocks per wavefront).  A
// 64x64 tile per wavefront (4x4 blocks) measured no faster on the GEMM
// workload with m, n, k ~ U{8..128}, at twice the registers.
constexpr int FAST_TS = 32;

// The complex instantiation is held to at least three wavefronts per SIMD:
// left alone it took 108 VGPRs and 64 AGPRs (two per SIMD); bounded it fits
// 128 registers with the same 24-byte spill and runs four (25.5 -> 24.4 ms
// for 200,000 products).
template <bool CPLX>
__global__ __launch_bounds__(64 * FW, (CPLX ? 3 : 1)) void k_gemm_mfma(MfhGemmBatch B, int count, GemmWs ws)
{
    constexpr int TS = FAST_TS;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const TileWalk walk = tile_walk(ws.tiles[count], FW, wave);
    for (int tile = walk.next; tile < walk.end; tile += walk.stride) {
        const int p = __builtin_amdgcn_readfirstlane(ws.map[tile]);
        const int local = tile - ws.tiles[p];
        const int m = B.m[p], n = B.n[p];
        const int mt = (m + TS - 1) / TS;
        const int i0 = (local % mt) * TS, j0 = (local / mt) * TS;
        // 16-row / 16-column blocks of this tile that hold any of C
        const int nbi = min(TS / 16, (m - i0 + 15) / 16), nbj = min(TS / 16, (n - j0 + 15) / 16);
        if (CPLX) {
            switch (nbi * 2 + nbj) {
            case 6: ztile<2, 2>(B, p, i0, j0, lane); break;
            case 5: ztile<2, 1>(B, p, i0, j0, lane); break;
            case 4: ztile<1, 2>(B, p, i0, j0, lane); break;
            default: ztile<1, 1>(B, p, i0, j0, lane); break;
            }
        } else {
            switch (nbi * 2 + nbj) {
            case 6: dtile<2, 2>(B, p, i0, j0, lane); break;
            case 5: dtile<2, 1>(B, p, i0, j0, lane); break;
            case 4: dtile<1, 2>(B, p, i0, j0, lane); break;
            default: dtile<1, 1>(B, p, i0, j0, lane); break;
            }
        }
    }
}

// ---- FAST mode, m, n <= 128: one workgroup per product ----------------------
// The tile kernel above has each wavefront read its own 32-row panel of op(A)
// and 32-column panel of op(B), so a product with mt x nt tiles reads its
// operands about 2 mt nt / (mt + nt) times from HBM (2.7x the algorithmic
// bytes measured on m, n, k ~ U{8..128}).  Here one workgroup of four
// wavefronts owns a whole product: chunks of PKC k-values of op(A) (m x PKC)
// and op(B) (PKC x n) are loaded once, in coalesced runs along whichever
// index is contiguous in memory, into LDS (k-major, as the MFMA fragments
// want them), and the wavefronts take the product's 32x32 tiles of C
// round-robin (at most PQ each), keeping their accumulators across
// chunks.  The next chunk's loads are in flight during the current chunk's
// MFMAs.  Values past k are stored as zero; LDS rows past m or n are never
// combined into a stored element.
constexpr int PW = 8;          // wavefronts per workgroup
constexpr int PKC = 16;        // k values per chunk
constexpr int PMAX = 128;      // largest m and n of this path
constexpr int PLD = PMAX + 4;  // LDS row length (doubles)
constexpr int PPER = PMAX * PKC / (64 * PW);  // chunk elements per thread and operand
constexpr int PQ = (PMAX / 32) * (PMAX / 32) / PW;  // tiles per wavefront

// Element e of a chunk, over a fixed frame of PMAX rows (shifts, no
// division): op(A)(r, kk) is read along r for NoTrans (A[r + kk lda]) and
// along kk for Trans (A[kk + r lda]); op(B)(kk, c) along kk for NoTrans
// (B[kk + c ldb]) and along c for Trans (B[c + kk ldb]).  Lanes outside the
// matrix or past k load nothing and store zero.
__device__ __forceinline__ void chunk_pos(bool kfast, int e, int& r, int& kk)
{
    if (kfast) {
        kk = e & (PKC - 1);
        r = e / PKC;
    } else {
        r = e & (PMAX - 1);
        kk = e / PMAX;
    }
}

__global__ __launch_bounds__(64 * PW) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_gemm_prod(
    MfhGemmBatch B, int count, const int* cnt)
{
    static_assert((PKC & (PKC - 1)) == 0 && (PMAX & (PMAX - 1)) == 0, "shift mapping");
    __shared__ double As[PKC * PLD];
    __shared__ double Bs[PKC * PLD];
    typedef const __attribute__((address_space(1))) double* GSrc;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int p = blockIdx.x; p < count; p += gridDim.x) {
        if (cnt[p] == 0) continue;  // invalid: info written by the check
        const int m = B.m[p], n = B.n[p], k = B.k[p];
        const bool ta = B.transa[p] == MFH_TRANS, tb = B.transb[p] == MFH_TRANS;
        const int lda = B.lda[p], ldb = B.ldb[p];
        const double* A = B.a[p];
        const double* Bm = B.b[p];
        const int mt = (m + 31) >> 5, nt = (n + 31) >> 5;
        const int T = mt * nt;
        double ra[PPER], rb[PPER];
        auto load = [&](int k0) {
#pragma unroll
            for (int u = 0; u < PPER; ++u) {
                const int e = tid + u * 64 * PW;
                int r, kk, c, kb;
                chunk_pos(ta, e, r, kk);
                chunk_pos(!tb, e, c, kb);
                const bool ina = r < m && k0 + kk < k;
                const bool inb = c < n && k0 + kb < k;
                ra[u] = ina ? ((GSrc)A)[ta ? (long)(k0 + kk) + (long)r * lda : (long)r + (long)(k0 + kk) * lda] : 0.0;
                rb[u] = inb ? ((GSrc)Bm)[tb ? (long)c + (long)(k0 + kb) * ldb : (long)(k0 + kb) + (long)c * ldb] : 0.0;
            }
        };
        mfh_d4 acc[PQ][2][2];
#pragma unroll
        for (int q = 0; q < PQ; ++q)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[q][i][j] = mfh_d4{0.0, 0.0, 0.0, 0.0};
        const int nchunks = (k + PKC - 1) / PKC;
        load(0);
        for (int ch = 0; ch < nchunks; ++ch) {
            __syncthreads();  // every wavefront is done with the previous chunk
#pragma unroll
            for (int u = 0; u < PPER; ++u) {
                const int e = tid + u * 64 * PW;
                int r, kk, c, kb;
                chunk_pos(ta, e, r, kk);
                chunk_pos(!tb, e, c, kb);
                As[kk * PLD + r] = ra[u];
                Bs[kb * PLD + c] = rb[u];
            }
            __syncthreads();
            if (ch + 1 < nchunks) load((ch + 1) * PKC);
#pragma unroll
            for (int q = 0; q < PQ; ++q) {
                const int t = wave + PW * q;
                if (t < T) {
                    const int i0 = (t % mt) * 32, j0 = (t / mt) * 32;
                    const bool i1 = i0 + 16 < m, j1 = j0 + 16 < n;
#pragma unroll
                    for (int ks = 0; ks < PKC / 4; ++ks) {
                        const int row = (ks * 4 + (lane >> 4)) * PLD + (lane & 15);
                        const double a0 = As[row + i0], a1 = As[row + i0 + 16];
                        const double b0 = Bs[row + j0], b1 = Bs[row + j0 + 16];
                        acc[q][0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[q][0][0], 0, 0, 0);
                        if (j1) acc[q][0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[q][0][1], 0, 0, 0);
                        if (i1) acc[q][1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[q][1][0], 0, 0, 0);
                        if (i1 && j1)
                            acc[q][1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[q][1][1], 0, 0, 0);
                    }
                }
            }
        }
        const double alpha = B.alpha[p], beta = B.beta[p];
        double* C = B.c[p];
        const long ldc = B.ldc[p];
#pragma unroll
        for (int q = 0; q < PQ; ++q) {
            const int t = wave + PW * q;
            if (t < T) {
                const int i0 = (t % mt) * 32, j0 = (t / mt) * 32;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int gi = i0 + 16 * i + (lane & 15);
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int gj = j0 + 16 * j + (lane >> 4) + 4 * r;
                            if (gi < m && gj < n) {
                                double* cp = C + gi + gj * ldc;
                                double v = alpha * acc[q][i][j][r];
                                if (beta != 0.0) v = __builtin_fma(beta, *cp, v);
                                *cp = v;
                            }
                        }
                }
            }
        }
    }
}

}  // namespace

// Workspace layout of one GEMM call (ints): per-product tile counts, chunk
// sums, then the tile -> product map.
struct GemmWsLayout {
    long maxtiles;
    int nchunks;
    size_t o_cnt, o_bsum, o_map, bytes;
};

static GemmWsLayout gemm_layout(int count, int max_m, int max_n, bool fast)
{
    const int ts = fast ? FAST_TS : TM;
    GemmWsLayout L;
    L.maxtiles = (long)count * ((max_m + ts - 1) / ts) * ((max_n + ts - 1) / ts);
    L.nchunks = (count + SCAN_CHUNK - 1) / SCAN_CHUNK;
    auto up = [](size_t v) { return (v + 63) / 64 * 64; };
    L.o_cnt = up((size_t)count + 1);
    L.o_bsum = L.o_cnt + up(count);
    L.o_map = L.o_bsum + up(L.nchunks);
    L.bytes = sizeof(int) * (L.o_map + (size_t)(L.maxtiles > 0 ? L.maxtiles : 0));
    return L;
}

// MFH_GEMM_TILES=1 in the environment keeps FAST FP64 products with m, n <=
// 128 on the tile kernel (for comparing the two).
static bool prod_kernel_off()
{
    static const bool off = [] {
        const char* e = getenv("MFH_GEMM_TILES");
        return e != nullptr && e[0] == '1';
    }();
    return off;
}

size_t mfh_gemm_ws_bytes(int count, int max_m, int max_n, bool fast)
{
    return gemm_layout(count, max_m, max_n, fast).bytes;
}

int mfh_launch_gemm(mfh_handle* h, int count, const MfhGemmBatch& B, int max_m, int max_n, int max_k, bool cplx)
{
    (void)max_k;
    const bool fast = h->math_mode == MFH_MATH_FAST;
    const int ts = fast ? FAST_TS : TM;
    const GemmWsLayout Lw = gemm_layout(count, max_m, max_n, fast);
    const long maxtiles = Lw.maxtiles;
    if (maxtiles > 0x7fffffffL - 64) return mfh_fail(MFH_ERR_UNSUPPORTED, "too many GEMM tiles in one call");
    const int nchunks = Lw.nchunks;
    const size_t o_cnt = Lw.o_cnt, o_bsum = Lw.o_bsum, o_map = Lw.o_map;
    int st = mfh_ws_reserve(h, Lw.bytes);
    if (st) return st;
    int* w = (int*)h->ws;
    GemmWs ws{w, w + o_cnt, w + o_bsum, w + o_map};
    hipLaunchKernelGGL(k_gemm_validate_scan, dim3(nchunks), dim3(SCAN_T), 0, h->stream, B, count, ws, max_m, max_n,
                       cplx ? 1 : 0, ts);
    MFH_HIP_CHECK(hipGetLastError());
    if (fast && !cplx && max_m <= PMAX && max_n <= PMAX && !prod_kernel_off()) {
        // every product fits one workgroup's LDS chunks: no tile list needed
        int per_cu = 0;
        MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gemm_prod, 64 * PW, 0));
        if (per_cu < 1) per_cu = 1;
        long grid = (long)per_cu * h->num_cus;
        if (grid > count) grid = count;
        hipLaunchKernelGGL(k_gemm_prod, dim3((unsigned)grid), dim3(64 * PW), 0, h->stream, B, count,
                           (const int*)ws.cnt);
        MFH_HIP_CHECK(hipGetLastError());
        return MFH_OK;
    }
    hipLaunchKernelGGL(k_gemm_map, dim3(nchunks), dim3(SCAN_T), 0, h->stream, ws, count, nchunks);
    MFH_HIP_CHECK(hipGetLastError());
    if (fast) {
        int per_cu = 0;
        if (cplx) MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gemm_mfma<true>, 64 * FW, 0));
        else MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gemm_mfma<false>, 64 * FW, 0));
        if (per_cu < 1) per_cu = 1;
        long grid = (long)per_cu * h->num_cus;
        const long groups = (maxtiles + FW - 1) / FW;
        if (grid > groups) grid = groups;
        grid = xcd_grid(grid);
        if (grid < 1) grid = 1;
        if (cplx) hipLaunchKernelGGL(k_gemm_mfma<true>, dim3((unsigned)grid), dim3(64 * FW), 0, h->stream, B, count, ws);
        else hipLaunchKernelGGL(k_gemm_mfma<false>, dim3((unsigned)grid), dim3(64 * FW), 0, h->stream, B, count, ws);
        MFH_HIP_CHECK(hipGetLastError());
        return MFH_OK;
    }
    int per_cu = 0;
    if (cplx) MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_zgemm_exact, GT, 0));
    else MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dgemm_exact, GT, 0));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > maxtiles) grid = maxtiles;
    grid = xcd_grid(grid);
    if (grid < 1) grid = 1;
    if (cplx) hipLaunchKernelGGL(k_zgemm_exact, dim3((unsigned)grid), dim3(GT), 0, h->stream, B, count, ws);
    else hipLaunchKernelGGL(k_dgemm_exact, dim3((unsigned)grid), dim3(GT), 0, h->stream, B, count, ws);
    MFH_HIP_CHECK(hipGetLastError());
    return MFH_OK;
}
