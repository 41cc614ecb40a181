// mfh_fast.hip -- FAST-mode batched FP64 Cholesky / SPD solve for gfx950.
//
// The exact kernels (mfh_exact.hip) reproduce MatFlat's operation order bit
// for bit, which forces one unfused multiply and one add per product and a
// dot-product (Crout) schedule whose parallelism shrinks to the rows below
// the current column.  FAST mode gives up bit-identity for throughput: the
// results match MatFlat to rounding (backward error of order n*eps) and are
// checked against the oracle by residual in the tests.
//
// Design (one wavefront per matrix, n <= 135 rows):
//  * right-looking Cholesky: at step k the column is finalised
//    (L[k,k] = sqrt(a_kk), L[i,k] = a_ik / L[k,k]) and the whole trailing
//    lower triangle takes the rank-1 update a_ij -= L[i,k] L[j,k] with FMA;
//  * the lower triangle lives in VGPRs, block-cyclic over an 8 x 8 lane
//    grid: lane (lr, lc) owns elements (lr + 8a, lc + 8b), a >= b, so every
//    update reads RB + CB column values from LDS for up to RB*(RB+1)/2 FMAs
//    and the lanes stay busy until the trailing triangle is a few blocks;
//  * LDS holds only the current column, the diagonal and the solve vector
//    (a few KB), so occupancy is set by VGPRs, not by the matrix size;
//  * the right-hand side of the SPD solve is carried as an extra row n of
//    the triangle (its "L[n,k]" is the forward substitution), and the
//    back substitution runs as a bottom-up column sweep from LDS;
//  * loads of the upper triangle and stores of L (and of the zeroed strict
//    upper triangle) touch each HBM byte once, eight contiguous doubles per
//    lane group.
#include "mfh_device.h"
#include "mfh_internal.h"


#pragma clang fp contract(fast)

namespace {

template <int RB>
struct Tri {
    static constexpr int size = RB * (RB + 1) / 2;
    static constexpr int idx(int a, int b) { return a * (a + 1) / 2 + b; }
};

// LDS per matrix (doubles): col[2][NR] diag[NR] rdiag[NR] y[NR] x[NR]
// junk[64] + flags.  Lanes that must not store write to junk[lane], so
// conditional stores stay branch-free.
__host__ __device__ constexpr int fast_lds_doubles(int nr) { return 6 * nr + 66; }

// Per-wave state of one factorization, shared by the block steps below.
struct FastCtx {
    double* colb;
    double* dgl;
    double* rdg;
    double* junk;
    int* failp;
    int lane, lr, lc, n, N, NR;
};

// Finalise column k, which lies in block column KB: the diagonal owner's
// value (lane (k%8, k%8)) is broadcast, the 8 lanes with lc == k%8 scale
// their column-k elements and publish them in colb[k & 1].  Returns false
// when the diagonal is not positive.
template <int RB, int KB>
__device__ __forceinline__ bool fast_finalize(double (&acc)[Tri<RB>::size], const FastCtx& c, int k)
{
    const int src = (k & 7) * 9;  // lane (k%8) + 8*(k%8)
    const double d = __shfl(acc[Tri<RB>::idx(KB, KB)], src, MFH_WAVE);
    if (!(d > 0)) {
        if (c.lane == 0) *c.failp = k + 1;
        return false;
    }
    const double lkk = __builtin_sqrt(d);
    const double rk = 1.0 / lkk;
    if (c.lane == 0) {
        c.dgl[k] = lkk;
        c.rdg[k] = rk;
    }
    double* ck = c.colb + (k & 1) * c.NR;
    const bool owner = c.lc == (k & 7);
#pragma unroll
    for (int a = KB; a < RB; ++a) {
        const int i = c.lr + 8 * a;
        const bool ok = owner && i > k && i < c.N;
        const double t = acc[Tri<RB>::idx(a, KB)] * rk;
        acc[Tri<RB>::idx(a, KB)] = ok ? t : acc[Tri<RB>::idx(a, KB)];
        *(ok ? ck + i : c.junk + c.lane) = t;
    }
    return true;
}

// Rank-1 update with column k (block column KB) of every block (a, bb),
// a >= bb >= KB, except block column SKIP (already updated by the look-ahead;
// SKIP = -1 for none).
template <int RB, int KB, int SKIP>
__device__ __forceinline__ void fast_update(double (&acc)[Tri<RB>::size], const double (&rv)[RB], const double (&cv)[RB])
{
#pragma unroll
    for (int a = KB; a < RB; ++a)
#pragma unroll
        for (int bb = KB; bb <= a; ++bb) {
            if (bb == SKIP) continue;
            acc[Tri<RB>::idx(a, bb)] -= rv[a] * cv[bb];
        }
}

template <int RB, int COL>
__device__ __forceinline__ void fast_update_col(double (&acc)[Tri<RB>::size], const double (&rv)[RB], const double (&cv)[RB])
{
#pragma unroll
    for (int a = COL; a < RB; ++a) acc[Tri<RB>::idx(a, COL)] -= rv[a] * cv[COL];
}

// Steps k = 8*KB .. min(8*KB + 8, n) - 1, then the next block column.
template <int RB, int KB>
__device__ __forceinline__ bool fast_block_steps(double (&acc)[Tri<RB>::size], const FastCtx& c)
{
    const int kend = (8 * KB + 8 < c.n) ? 8 * KB + 8 : c.n;
    for (int k = 8 * KB; k < kend; ++k) {
        const double* ck = c.colb + (k & 1) * c.NR;
        double rv[RB], cv[RB];
#pragma unroll
        for (int a = 0; a < RB; ++a) {
            if (a < KB) {
                rv[a] = 0.0;
                cv[a] = 0.0;
                continue;
            }
            // Rows need no mask: an element (i, j) with i <= k and j > k is
            // in the (unused) upper part of a diagonal block, and rows i >= N
            // are never stored.  Columns j <= k are finished and must not
            // change, so their factor is zero.
            const int i = c.lr + 8 * a;
            rv[a] = ck[i < c.N ? i : c.N - 1];
            const int j = c.lc + 8 * a;
            const double y = ck[j < c.N ? j : c.N - 1];
            cv[a] = (j > k) ? y : 0.0;
        }
        bool ok = true;
        // Look-ahead: finish column k+1 first (its block column is KB, or
        // KB+1 at the last step of the block), then the rest of the update.
        if (k + 1 < c.n) {
            if ((k & 7) != 7) {
                fast_update_col<RB, KB>(acc, rv, cv);
                ok = fast_finalize<RB, KB>(acc, c, k + 1);
                fast_update<RB, KB, KB>(acc, rv, cv);
            } else {
                if constexpr (KB + 1 < RB) {
                    fast_update_col<RB, KB + 1>(acc, rv, cv);
                    ok = fast_finalize<RB, KB + 1>(acc, c, k + 1);
                    fast_update<RB, KB, KB + 1>(acc, rv, cv);
                }
            }
        } else {
            fast_update<RB, KB, -1>(acc, rv, cv);
        }
        __syncthreads();
        if (!ok) return false;
    }
    if constexpr (KB + 1 < RB) {
        if (8 * KB + 8 < c.n) return fast_block_steps<RB, KB + 1>(acc, c);
    }
    return true;
}

// Back substitution steps j = 8*KB+7 .. 8*KB (descending), then block KB-1.
template <int RB, int KB>
__device__ __forceinline__ void fast_back_steps(const double (&acc)[Tri<RB>::size], const FastCtx& c, double* yv,
                                                double* xv)
{
    const int jhi = (8 * KB + 7 < c.n - 1) ? 8 * KB + 7 : c.n - 1;
    for (int j = jhi; j >= 8 * KB; --j) {
        const double xj = yv[j] * c.rdg[j];
        if (c.lane == 0) xv[j] = xj;
        const bool mine = c.lr == (j & 7);
#pragma unroll
        for (int bb = 0; bb <= KB; ++bb) {
            const int r = c.lc + 8 * bb;
            const bool ok = mine && r < j;
            const int rc = r < j ? r : 0;
            const double nv = yv[rc] - acc[Tri<RB>::idx(KB, bb)] * xj;
            *(ok ? yv + r : c.junk + c.lane) = nv;
        }
        __syncthreads();
    }
    if constexpr (KB > 0) fast_back_steps<RB, KB - 1>(acc, c, yv, xv);
}

template <int RB, int KB>
__device__ __forceinline__ void fast_back_from(const double (&acc)[Tri<RB>::size], const FastCtx& c, double* yv,
                                               double* xv, int nb)
{
    // start at the block that holds row n-1
    if (nb == KB) {
        fast_back_steps<RB, KB>(acc, c, yv, xv);
        return;
    }
    if constexpr (KB > 0) fast_back_from<RB, KB - 1>(acc, c, yv, xv, nb);
}

template <int RB, bool POSV>
__global__ __launch_bounds__(64) void k_chol_fast(MfhCholBatch B, MfhQueue q, int NR)
{
    extern __shared__ double lds[];
    FastCtx c;
    c.colb = lds;  // column buffers, double-buffered by step parity
    c.dgl = lds + 2 * NR;
    c.rdg = lds + 3 * NR;
    double* yv = lds + 4 * NR;
    double* xv = lds + 5 * NR;
    c.junk = lds + 6 * NR;
    c.failp = (int*)(lds + 6 * NR + 64);
    c.lane = threadIdx.x;
    c.lr = c.lane & 7;
    c.lc = c.lane >> 3;
    c.NR = NR;
    const int lr = c.lr, lc = c.lc, lane = c.lane;
    constexpr int T = Tri<RB>::size;

    for (int it = blockIdx.x;; it += gridDim.x) {
        // One matrix per wavefront: its id and sizes are wave-uniform, which
        // keeps every step and block test below a scalar branch.
        const int b = __builtin_amdgcn_readfirstlane(mfh_item(q, it));
        if (b < 0) break;
        const int n = __builtin_amdgcn_readfirstlane(B.n[b]);
        const int N = POSV ? n + 1 : n;  // rows of the triangle
        const int lda = __builtin_amdgcn_readfirstlane(B.lda[b]);
        c.n = n;
        c.N = N;
        double* A = B.a[b];
        const double* X = POSV ? B.x[b] : nullptr;

        // Load element (i, j), i >= j: the upper-triangle entry A[j + i*lda];
        // row n of the SPD solve holds b.  Addresses are clamped into the
        // matrix and the value selected, so the loads are unconditional.
        double acc[T];
#pragma unroll
        for (int a = 0; a < RB; ++a) {
            const int i = lr + 8 * a;
            const int ic = i < n ? i : n - 1;
#pragma unroll
            for (int bb = 0; bb <= a; ++bb) {
                const int j = lc + 8 * bb;
                const int jc = j < ic ? j : ic;
                double v = A[jc + (size_t)ic * lda];
                v = (i < n && j <= i) ? v : 0.0;
                if (POSV && a == (n >> 3)) {
                    const double xb = X[j < n ? j : n - 1];
                    v = (i == n && j < n) ? xb : v;
                }
                acc[Tri<RB>::idx(a, bb)] = v;
            }
        }
        if (lane == 0) *c.failp = 0;
        bool ok = fast_finalize<RB, 0>(acc, c, 0);
        __syncthreads();
        if (ok) ok = fast_block_steps<RB, 0>(acc, c);
        const int fail = __builtin_amdgcn_readfirstlane(*(volatile int*)c.failp);

        if (fail) {
            if (lane == 0) B.info[b] = fail;
        } else {
            // Store L, the diagonal and the zeroed strict upper triangle: in
            // a diagonal block each lane writes its own (i, j); off-diagonal
            // blocks write L at (i, j) and 0 at (j, i).
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                const int i = lr + 8 * a;
#pragma unroll
                for (int bb = 0; bb <= a; ++bb) {
                    const int j = lc + 8 * bb;
                    const double v = acc[Tri<RB>::idx(a, bb)];
                    if (bb == a) {
                        if (i < n && j < n) A[i + (size_t)j * lda] = (i > j) ? v : (i == j ? c.dgl[i] : 0.0);
                    } else if (i < n) {
                        A[i + (size_t)j * lda] = v;
                        A[j + (size_t)i * lda] = 0.0;
                    }
                }
            }
            if (lane == 0) B.info[b] = 0;
            if (POSV) {
                // y = row n of the triangle, then the Lower/Trans sweep
                // (Blas.SolveTriangular.cs:224-231 order) with x in xv.
#pragma unroll
                for (int a = 0; a < RB; ++a) {
                    if (a != (n >> 3)) continue;
#pragma unroll
                    for (int bb = 0; bb <= a; ++bb) {
                        const int j = lc + 8 * bb;
                        const bool okw = lr == (n & 7) && j < n;
                        *(okw ? yv + j : c.junk + lane) = acc[Tri<RB>::idx(a, bb)];
                    }
                }
                __syncthreads();
                fast_back_from<RB, RB - 1>(acc, c, yv, xv, (n - 1) >> 3);
                double* Xo = B.x[b];
                for (int r = lane; r < n; r += 64) Xo[r] = xv[r];
            }
        }
        __syncthreads();
    }
}

template <int RB, bool POSV>
int launch_fast(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int nr, long count)
{
    const size_t lds = (size_t)fast_lds_doubles(nr) * 8;
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chol_fast<RB, POSV>, 64, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > count) grid = count;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL((k_chol_fast<RB, POSV>), dim3((unsigned)grid), dim3(64), lds, h->stream, B, q, nr);
    MFH_HIP_CHECK(hipGetLastError());
    return MFH_OK;
}

template <bool POSV>
int dispatch_fast(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int nr, long count)
{
    const int rb = (nr + 7) / 8;
    switch (rb) {
    case 1: return launch_fast<1, POSV>(h, B, q, nr, count);
    case 2: return launch_fast<2, POSV>(h, B, q, nr, count);
    case 3: return launch_fast<3, POSV>(h, B, q, nr, count);
    case 4: return launch_fast<4, POSV>(h, B, q, nr, count);
    case 5: return launch_fast<5, POSV>(h, B, q, nr, count);
    case 6: return launch_fast<6, POSV>(h, B, q, nr, count);
    case 7: return launch_fast<7, POSV>(h, B, q, nr, count);
    case 8: return launch_fast<8, POSV>(h, B, q, nr, count);
    case 9: return launch_fast<9, POSV>(h, B, q, nr, count);
    case 10: return launch_fast<10, POSV>(h, B, q, nr, count);
    case 11: return launch_fast<11, POSV>(h, B, q, nr, count);
    case 12: return launch_fast<12, POSV>(h, B, q, nr, count);
    case 13: return launch_fast<13, POSV>(h, B, q, nr, count);
    case 14: return launch_fast<14, POSV>(h, B, q, nr, count);
    case 15: return launch_fast<15, POSV>(h, B, q, nr, count);
    case 16: return launch_fast<16, POSV>(h, B, q, nr, count);
    case 17: return launch_fast<17, POSV>(h, B, q, nr, count);
    default: return mfh_fail(MFH_ERR_UNSUPPORTED, "fast Cholesky: too many rows");
    }
}

}  // namespace

int mfh_launch_chol_fast_bucket(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int nr, long count, bool posv)
{
    return posv ? dispatch_fast<true>(h, B, q, nr, count) : dispatch_fast<false>(h, B, q, nr, count);
}
