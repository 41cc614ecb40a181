// mfh_fast.hip -- FAST-mode batched FP64 Cholesky / SPD solve for gfx950.
//
// The exact kernels (mfh_exact.hip) reproduce MatFlat's operation order bit
// for bit, which forces one unfused multiply and one add per product and a
// dot-product (Crout) schedule whose parallelism shrinks to the rows below
// the current column.  FAST mode gives up bit-identity for throughput: the
// results match MatFlat to rounding (backward error of order n*eps) and are
// checked against the oracle by residual in the tests.
//
// Design (one wavefront per matrix, at most 8*RB rows, RB <= 17):
//  * right-looking Cholesky: at step k the whole trailing lower triangle
//    takes the rank-1 update a_ij -= L[i,k] L[j,k] with FMA, and column k+1
//    is finalised first (look-ahead) so its square root and its publication
//    overlap the rest of the update;
//  * the lower triangle lives in VGPRs, block-cyclic over an 8 x 8 lane
//    grid: lane (lr, lc) owns elements (lr + 8a, lc + 8b), a >= b.  The
//    accumulators keep column k unscaled (a_ik, not a_ik / L[k,k]); only
//    the published copy is scaled, and one pass at the end scales the
//    registers, so a step carries no per-element selects;
//  * a finished column is published in LDS in a permuted order (element i
//    at (i % 8) * RBP + i / 8), so each lane reads the RB row values and RB
//    column values it needs as 16-byte pairs, double-buffered by step
//    parity so a step needs one barrier;
//  * the diagonal element is broadcast with readlane (k is wave-uniform)
//    and 1/sqrt comes from the hardware estimate plus two Newton steps;
//  * the right-hand side of the SPD solve is carried as an extra row n of
//    the triangle (its "L[n,k]" is the forward substitution), and the back
//    substitution runs as a bottom-up column sweep through LDS;
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

// Row stride of a permuted column buffer (even: pairs stay 16-byte aligned).
template <int RB>
constexpr int rbp() { return (RB + 1) & ~1; }

// LDS per workgroup (doubles): col[2][8 * RBP], dgl[8RB], rdg[8RB],
// yv[8RB], xv[8RB], one flag.
template <int RB>
constexpr int fast_lds_doubles() { return 16 * rbp<RB>() + 32 * RB + 2; }

// Global-memory pointer: loads and stores through it use the global (not
// flat) instructions, with a wave-uniform base and 32-bit lane offsets.
typedef __attribute__((address_space(1))) double* GPtr;

struct FastCtx {
    double* col;  // two permuted column buffers
    double* dgl;  // L[k,k]
    double* rdg;  // 1 / L[k,k]
    int* failp;
    int lane, lr, lc, n;
};

// 1/sqrt(d) for a positive normal d: hardware estimate, then two Newton
// steps y <- y + y (1 - d y^2) / 2 (each squares the relative error).
__device__ __forceinline__ double fast_rsqrt(double d)
{
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double e = __builtin_fma(-(d * y), y, 1.0);
        y = __builtin_fma(0.5 * y, e, y);
    }
    return y;
}

// v[a] for a >= KB from row r of a permuted column buffer.
template <int RB, int KB>
__device__ __forceinline__ void load_col(const double* cp, int r, double (&v)[RB])
{
    const double2* p = reinterpret_cast<const double2*>(cp + r * rbp<RB>());
#pragma unroll
    for (int a = KB & ~1; a < RB; a += 2) {
        const double2 t = p[a / 2];
        if (a >= KB) v[a] = t.x;
        if (a + 1 < RB) v[a + 1] = t.y;
    }
}

// Finalise column k (block column KB): broadcast the diagonal a_kk from
// lane (k%8, k%8), record L[k,k] and its reciprocal, and let the 8 lanes
// with lc == k%8 publish L[i,k] = a_ik / L[k,k] for their rows into cn.
// Returns false when a_kk is not positive.
template <int RB, int KB>
__device__ __forceinline__ bool fast_finalize(const double (&acc)[Tri<RB>::size], const FastCtx& c, int k,
                                              double* cn)
{
    const int src = (k & 7) * 9;
    const double dv = acc[Tri<RB>::idx(KB, KB)];
    const int lo = __builtin_amdgcn_readlane(__double2loint(dv), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(dv), src);
    const double d = __hiloint2double(hi, lo);
    if (!(d > 0.0)) {
        if (c.lane == 0) *c.failp = k + 1;
        return false;
    }
    const double y = fast_rsqrt(d);
    if (c.lane == 0) {
        c.dgl[k] = d * y;
        c.rdg[k] = y;
    }
    if (c.lc == (k & 7)) {
        double2* p = reinterpret_cast<double2*>(cn + c.lr * rbp<RB>());
#pragma unroll
        for (int a = KB & ~1; a < RB; a += 2) {
            const double x0 = (a >= KB) ? acc[Tri<RB>::idx(a >= KB ? a : KB, KB)] * y : 0.0;
            const double x1 = (a + 1 < RB) ? acc[Tri<RB>::idx(a + 1 < RB ? a + 1 : KB, KB)] * y : 0.0;
            p[a / 2] = make_double2(x0, x1);
        }
    }
    return true;
}

// Rank-1 update of every block (a, bb), a >= bb >= KB, except block column
// SKIP (already updated for the look-ahead; -1 for none).
template <int RB, int KB, int SKIP>
__device__ __forceinline__ void fast_update(double (&acc)[Tri<RB>::size], const double (&rv)[RB], const double (&cv)[RB])
{
#pragma unroll
    for (int bb = KB; bb < RB; ++bb) {
        if (bb == SKIP) continue;
#pragma unroll
        for (int a = bb; a < RB; ++a) acc[Tri<RB>::idx(a, bb)] -= rv[a] * cv[bb];
    }
}

template <int RB, int COL>
__device__ __forceinline__ void fast_update_col(double (&acc)[Tri<RB>::size], const double (&rv)[RB], const double (&cv)[RB])
{
#pragma unroll
    for (int a = COL; a < RB; ++a) acc[Tri<RB>::idx(a, COL)] -= rv[a] * cv[COL];
}

// Steps k = 8*KB .. min(8*KB + 8, n) - 1, then the next block column.
// Column k was published into buffer k & 1 by the previous step.
template <int RB, int KB>
__device__ __forceinline__ bool fast_block_steps(double (&acc)[Tri<RB>::size], const FastCtx& c)
{
    constexpr int CB = 8 * rbp<RB>();
    const int kend = (8 * KB + 8 < c.n) ? 8 * KB + 8 : c.n;
    for (int k = 8 * KB; k < kend; ++k) {
        const double* cp = c.col + (k & 1) * CB;
        double* cn = c.col + ((k + 1) & 1) * CB;
        double rv[RB], cv[RB];
        load_col<RB, KB>(cp, c.lr, rv);
        load_col<RB, KB>(cp, c.lc, cv);
        // Rows need no mask: an element (i, j) with i <= k < j lies in the
        // unused upper part of a diagonal block, and rows beyond the matrix
        // are never stored.  Finished columns j <= k must not change.
        if (c.lc + 8 * KB <= k) cv[KB] = 0.0;
        bool ok = true;
        if (k + 1 < c.n) {
            if ((k & 7) != 7) {
                fast_update_col<RB, KB>(acc, rv, cv);
                ok = fast_finalize<RB, KB>(acc, c, k + 1, cn);
                fast_update<RB, KB, KB>(acc, rv, cv);
            } else {
                if constexpr (KB + 1 < RB) {
                    fast_update_col<RB, KB + 1>(acc, rv, cv);
                    ok = fast_finalize<RB, KB + 1>(acc, c, k + 1, cn);
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

// Back substitution L^T x = y, columns j = min(8*KB+7, n-1) .. 8*KB
// (descending), then block KB-1: x_j = y_j / L[j,j], then y_r -= L[j,r] x_j
// for r < j by the 8 lanes that hold row j.
template <int RB, int KB>
__device__ __forceinline__ void fast_back_steps(const double (&acc)[Tri<RB>::size], const FastCtx& c, double* yv,
                                                double* xv)
{
    const int jhi = (8 * KB + 7 < c.n - 1) ? 8 * KB + 7 : c.n - 1;
    for (int j = jhi; j >= 8 * KB; --j) {
        const double xj = yv[j] * c.rdg[j];
        if (c.lane == 0) xv[j] = xj;
        if (c.lr == (j & 7)) {
#pragma unroll
            for (int bb = 0; bb <= KB; ++bb) {
                const int r = c.lc + 8 * bb;
                if (r < j) yv[r] -= acc[Tri<RB>::idx(KB, bb)] * xj;
            }
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
__global__ __launch_bounds__(64) void k_chol_fast(MfhCholBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    constexpr int NR = 8 * RB;
    constexpr int T = Tri<RB>::size;
    FastCtx c;
    c.col = lds;
    c.dgl = lds + 16 * rbp<RB>();
    c.rdg = c.dgl + NR;
    double* yv = c.rdg + NR;
    double* xv = yv + NR;
    c.failp = (int*)(xv + NR);
    c.lane = threadIdx.x;
    c.lr = c.lane & 7;
    c.lc = c.lane >> 3;
    const int lr = c.lr, lc = c.lc, lane = c.lane;

    for (int it = blockIdx.x;; it += gridDim.x) {
        // One matrix per wavefront: its id and sizes are wave-uniform, which
        // keeps every step and block test below a scalar branch.
        const int b = __builtin_amdgcn_readfirstlane(mfh_item(q, it));
        if (b < 0) break;
        const int n = __builtin_amdgcn_readfirstlane(B.n[b]);
        const int lda = __builtin_amdgcn_readfirstlane(B.lda[b]);
        c.n = n;
        GPtr A = (GPtr)B.a[b];

        // Load element (i, j), i >= j: the upper-triangle entry A[j + i*lda].
        // Offsets are clamped into the matrix so every load is unconditional
        // (32-bit offsets from a wave-uniform base), and the value selected.
        double acc[T];
#pragma unroll
        for (int a = 0; a < RB; ++a) {
            const int i = lr + 8 * a;
            const int ic = i < n ? i : n - 1;
#pragma unroll
            for (int bb = 0; bb <= a; ++bb) {
                const int j = lc + 8 * bb;
                const int jc = j < ic ? j : ic;
                const double v = A[(unsigned)(jc + ic * lda)];
                acc[Tri<RB>::idx(a, bb)] = (i < n && j <= i) ? v : 0.0;
            }
        }
        if (POSV) {
            // Row n of the triangle holds b (lanes with lr == n % 8, block n / 8).
            GPtr X = (GPtr)B.x[b];
            double xb[RB];
#pragma unroll
            for (int bb = 0; bb < RB; ++bb) {
                const int j = lc + 8 * bb;
                xb[bb] = X[(unsigned)(j < n ? j : n - 1)];
            }
            const bool rown = lr == (n & 7);
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                if (a != (n >> 3)) continue;
#pragma unroll
                for (int bb = 0; bb <= a; ++bb) {
                    const int j = lc + 8 * bb;
                    acc[Tri<RB>::idx(a, bb)] = (rown && j < n) ? xb[bb] : acc[Tri<RB>::idx(a, bb)];
                }
            }
        }
        if (lane == 0) *c.failp = 0;
        bool ok = fast_finalize<RB, 0>(acc, c, 0, c.col);
        __syncthreads();
        if (ok) ok = fast_block_steps<RB, 0>(acc, c);
        const int fail = __builtin_amdgcn_readfirstlane(*(volatile int*)c.failp);

        if (fail) {
            // FAST mode leaves a matrix that is not positive definite unchanged.
            if (lane == 0) B.info[b] = fail;
        } else {
            // Scale the columns: a_ij / L[j,j] (the diagonal is stored from dgl).
#pragma unroll
            for (int bb = 0; bb < RB; ++bb) {
                const double r = c.rdg[lc + 8 * bb];
#pragma unroll
                for (int a = bb; a < RB; ++a) acc[Tri<RB>::idx(a, bb)] *= r;
            }
            // Store L, the diagonal and the zeroed strict upper triangle: in
            // a diagonal block each lane writes its own (i, j); off-diagonal
            // blocks write L at (i, j) and 0 at (j, i).  One exec-masked
            // region per block row.
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                const int i = lr + 8 * a;
                if (i < n) {
#pragma unroll
                    for (int bb = 0; bb < a; ++bb) {
                        const int j = lc + 8 * bb;
                        A[(unsigned)(i + j * lda)] = acc[Tri<RB>::idx(a, bb)];
                        A[(unsigned)(j + i * lda)] = 0.0;
                    }
                    const int j = lc + 8 * a;
                    if (j < n) {
                        const double v = acc[Tri<RB>::idx(a, a)];
                        A[(unsigned)(i + j * lda)] = (i > j) ? v : (i == j ? c.dgl[i] : 0.0);
                    }
                }
            }
            if (lane == 0) B.info[b] = 0;
            if (POSV) {
                // y = row n of the triangle, then the Lower/Trans sweep
                // (Blas.SolveTriangular.cs:224-231 order) with x in xv.
                if (lr == (n & 7)) {
#pragma unroll
                    for (int a = 0; a < RB; ++a) {
                        if (a != (n >> 3)) continue;
#pragma unroll
                        for (int bb = 0; bb <= a; ++bb) {
                            const int j = lc + 8 * bb;
                            if (j < n) yv[j] = acc[Tri<RB>::idx(a, bb)];
                        }
                    }
                }
                __syncthreads();
                fast_back_from<RB, RB - 1>(acc, c, yv, xv, (n - 1) >> 3);
                GPtr Xo = (GPtr)B.x[b];
                for (int r = lane; r < n; r += 64) Xo[r] = xv[r];
            }
        }
        __syncthreads();
    }
}

template <int RB, bool POSV>
int launch_fast(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, long count)
{
    const size_t lds = (size_t)fast_lds_doubles<RB>() * 8;
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chol_fast<RB, POSV>, 64, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > count) grid = count;
    if (grid < 1) grid = 1;
    int st = mfh_timing_begin(h, MFH_KERNEL_CHOL_FAST, 8 * RB);
    if (st) return st;
    hipLaunchKernelGGL((k_chol_fast<RB, POSV>), dim3((unsigned)grid), dim3(64), lds, h->stream, B, q);
    MFH_HIP_CHECK(hipGetLastError());
    return mfh_timing_end(h);
}

template <bool POSV>
int dispatch_fast(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count)
{
    switch (rb) {
    case 1: return launch_fast<1, POSV>(h, B, q, count);
    case 2: return launch_fast<2, POSV>(h, B, q, count);
    case 3: return launch_fast<3, POSV>(h, B, q, count);
    case 4: return launch_fast<4, POSV>(h, B, q, count);
    case 5: return launch_fast<5, POSV>(h, B, q, count);
    case 6: return launch_fast<6, POSV>(h, B, q, count);
    case 7: return launch_fast<7, POSV>(h, B, q, count);
    case 8: return launch_fast<8, POSV>(h, B, q, count);
    case 9: return launch_fast<9, POSV>(h, B, q, count);
    case 10: return launch_fast<10, POSV>(h, B, q, count);
    case 11: return launch_fast<11, POSV>(h, B, q, count);
    case 12: return launch_fast<12, POSV>(h, B, q, count);
    case 13: return launch_fast<13, POSV>(h, B, q, count);
    case 14: return launch_fast<14, POSV>(h, B, q, count);
    case 15: return launch_fast<15, POSV>(h, B, q, count);
    case 16: return launch_fast<16, POSV>(h, B, q, count);
    case 17: return launch_fast<17, POSV>(h, B, q, count);
    default: return mfh_fail(MFH_ERR_UNSUPPORTED, "fast Cholesky: too many rows");
    }
}

}  // namespace

int mfh_launch_chol_fast_bucket(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count, bool posv)
{
    return posv ? dispatch_fast<true>(h, B, q, rb, count) : dispatch_fast<false>(h, B, q, rb, count);
}
