// mfh_fast_kernels.h -- FAST-mode batched FP64 Cholesky / SPD solve for gfx950.
//
// The kernel templates are included by two translation units that build in
// parallel: mfh_fast.hip (Cholesky alone, and the bucket dispatcher) and
// mfh_fast_posv.hip (the SPD solve).
//
// The exact kernels (mfh_exact.hip) reproduce MatFlat's operation order bit
// for bit, which forces one unfused multiply and one add per product and a
// dot-product (Crout) schedule whose parallelism shrinks to the rows below
// the current column.  FAST mode gives up bit-identity for throughput: the
// results match MatFlat to rounding (backward error of order n*eps) and are
// checked against the oracle by residual in the tests.
//
// Design (one workgroup of NW = 1 or 2 wavefronts per matrix, at most 8*RB
// rows, RB <= 17):
//  * right-looking Cholesky on the unscaled (LDL^T-like) trailing matrix:
//    at step k every trailing element takes a_ij -= a_ik a_jk / a_kk with
//    FMA; column k+1 is updated and published first (look-ahead) so its
//    publication overlaps the rest of the update.  The step's critical path
//    holds one reciprocal, no square root: L[i,k] = a_ik / sqrt(a_kk) is
//    applied once, in a final pass over the registers;
//  * the lower triangle lives in VGPRs, block-cyclic over an 8 x 8 lane
//    grid: lane (lr, lc) of wave w owns elements (lr + 8a, lc + 8b) for
//    block rows a = w (mod NW) and b <= a.  Two wavefronts per matrix halve
//    the registers per lane (four were measured 2x slower: every step then
//    waits on a barrier across four wavefronts);
//  * a published column sits in LDS in a permuted order (element i of block
//    row a at [i % 8][(a % NW) * HALF + a / NW]), so a lane reads the row
//    values of its wave's block rows and the column values it needs as
//    16-byte pairs, double-buffered by step parity: one barrier per step;
//  * the right-hand side of the SPD solve is carried as an extra row n of
//    the triangle (its scaled "L[n,k]" is the forward substitution), and the
//    back substitution runs as a bottom-up column sweep through LDS;
//  * loads of the upper triangle and stores of L (and of the zeroed strict
//    upper triangle) touch each HBM byte once, eight contiguous doubles per
//    lane group, through 32-bit offsets from a wave-uniform base.
#pragma once
#include "mfh_device.h"
#include "mfh_internal.h"

#pragma clang fp contract(fast)

namespace {

// Global-memory pointer: loads and stores through it use the global (not
// flat) instructions.
typedef __attribute__((address_space(1))) double* GPtr;
typedef __attribute__((address_space(1))) char* GBytes;

// Element off of a matrix through a 32-bit byte offset from its base (the
// host checks lda * n < 2^28), so the address needs one VGPR.
__device__ __forceinline__ double gload(GPtr base, int off) { return *(GPtr)((GBytes)base + (unsigned)(off << 3)); }
__device__ __forceinline__ void gstore(GPtr base, int off, double v) { *(GPtr)((GBytes)base + (unsigned)(off << 3)) = v; }

// Block rows a = W, W + NW, ... of an RB-block triangle held by wave W, and
// the register index of block (m-th owned row, column block bb).
template <int RB, int NW, int W>
struct Own {
    static constexpr int na = (RB - W + NW - 1) / NW;
    __host__ __device__ __forceinline__ static constexpr int arow(int m) { return W + NW * m; }
    // sum of arow(t) + 1 over t < m, in closed form
    __host__ __device__ __forceinline__ static constexpr int base(int m)
    {
        return m * (W + 1) + NW * (m * (m - 1) / 2);
    }
    static constexpr int size = base(na) > 0 ? base(na) : 1;
    __host__ __device__ __forceinline__ static constexpr int idx(int m, int bb) { return base(m) + bb; }
    // first owned row index with block row >= kb
    __host__ __device__ __forceinline__ static constexpr int first(int kb)
    {
        return kb <= W ? 0 : (kb - W + NW - 1) / NW;
    }
};

// Register tile of wave W, and one value per owned block row.
template <int RB, int NW, int W>
using AccT = double[Own<RB, NW, W>::size];
template <int RB, int NW, int W>
using RowT = double[Own<RB, NW, W>::na > 0 ? Own<RB, NW, W>::na : 1];

// Column buffer geometry: HALF slots per wave (even), RS per lane row.
template <int RB, int NW>
__host__ __device__ __forceinline__ constexpr int half_slots() { return (((RB + NW - 1) / NW) + 1) & ~1; }
template <int RB, int NW>
__host__ __device__ __forceinline__ constexpr int row_stride() { return NW * half_slots<RB, NW>(); }
template <int RB, int NW>
__host__ __device__ __forceinline__ constexpr int slot_of(int a) { return (a % NW) * half_slots<RB, NW>() + a / NW; }

// LDS per workgroup (doubles): col[2][8 * RS], dgl[8RB], rdg[8RB], yv[8RB],
// xv[8RB].
template <int RB, int NW>
__host__ __device__ constexpr int fast_lds_doubles() { return 16 * row_stride<RB, NW>() + 32 * RB; }

struct FastCtx {
    double* col;  // two permuted column buffers
    double* dgl;  // L[k,k]
    double* rdg;  // 1 / L[k,k]
    int tid, lane, lr, lc, n;
};

// 1/d for a positive normal d: hardware estimate and two Newton steps.
__device__ __forceinline__ double fast_rcp(double d)
{
    double r = __builtin_amdgcn_rcp(d);
#pragma unroll
    for (int it = 0; it < 2; ++it) r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    return r;
}

// 1/sqrt(d) for a positive normal d: hardware estimate, then two Newton
// steps y <- y + y (1 - d y^2) / 2.
__device__ __forceinline__ double fast_rsqrt(double d)
{
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int it = 0; it < 2; ++it) y = __builtin_fma(0.5 * y, __builtin_fma(-(d * y), y, 1.0), y);
    return y;
}

// Publish column col (block column KB) of wave W's rows, unscaled: the lanes
// with lc == col % 8 write every slot of their wave's half (0 for rows above
// block row KB), so no stale value is ever read.
template <int RB, int NW, int W, int KB>
__device__ __forceinline__ void fast_publish(const AccT<RB, NW, W>& acc, const FastCtx& c, int col,
                                             double* cn)
{
    using O = Own<RB, NW, W>;
    constexpr int H = half_slots<RB, NW>();
    if (c.lc == (col & 7)) {
        double2* p = reinterpret_cast<double2*>(cn + c.lr * row_stride<RB, NW>() + W * H);
#pragma unroll
        for (int m = 0; m < H; m += 2) {
            const bool v0 = m < O::na && O::arow(m) >= KB;
            const bool v1 = m + 1 < O::na && O::arow(m + 1) >= KB;
            const double x0 = v0 ? acc[O::idx(v0 ? m : 0, v0 ? KB : 0)] : 0.0;
            const double x1 = v1 ? acc[O::idx(v1 ? m + 1 : 0, v1 ? KB : 0)] : 0.0;
            p[m / 2] = make_double2(x0, x1);
        }
    }
}

// Rank-1 update of wave W's blocks (m, bb), bb >= KB0: acc -= rs[m] * cv[bb],
// with rs the row values already divided by a_kk.
template <int RB, int NW, int W, int KB0>
__device__ __forceinline__ void fast_update(AccT<RB, NW, W>& acc, const RowT<RB, NW, W>& rs, const double (&cv)[RB])
{
    using O = Own<RB, NW, W>;
    // constant trip counts (the bounds are resolved after unrolling)
#pragma unroll
    for (int m = O::first(KB0); m < O::na; ++m)
#pragma unroll
        for (int bb = KB0; bb < RB; ++bb) {
            if (bb > O::arow(m)) continue;
            acc[O::idx(m, bb)] -= rs[m] * cv[bb];
        }
}

template <int RB, int NW, int W, int COL>
__device__ __forceinline__ void fast_update_col(AccT<RB, NW, W>& acc,
                                                const RowT<RB, NW, W>& rs, const double (&cv)[RB])
{
    using O = Own<RB, NW, W>;
#pragma unroll
    for (int m = O::first(COL); m < O::na; ++m) acc[O::idx(m, COL)] -= rs[m] * cv[COL];
}

// Steps k = 8*KB .. min(8*KB + 8, n) - 1, then the next block column.  Step
// k reads column k (unscaled, with a_kk at row k) from buffer k & 1, updates
// the column k+1 belongs to (block column KB, or KB+1 at the block's last
// step) and publishes it into the other buffer, then updates the rest of
// the trailing matrix.  Both candidate look-ahead block columns are updated
// before the (uniform) branch, so the bulk update is one straight path.  At
// the very last step the look-ahead publishes a column nobody reads.  The
// first pivot that is not positive is recorded in fail (the factorization
// goes on with garbage, which the caller discards).
template <int RB, int NW, int W, int KB>
__device__ __forceinline__ void fast_block_steps(AccT<RB, NW, W>& acc, const FastCtx& c, int& fail)
{
    using O = Own<RB, NW, W>;
    constexpr int H = half_slots<RB, NW>();
    constexpr int RS = row_stride<RB, NW>();
    constexpr int CB = 8 * RS;
    constexpr int M0 = O::first(KB);
    const int kend = (8 * KB + 8 < c.n) ? 8 * KB + 8 : c.n;
    for (int k = 8 * KB; k < kend; ++k) {
        const double* cp = c.col + (k & 1) * CB;
        double* cn = c.col + ((k + 1) & 1) * CB;
        const double d = cp[(k & 7) * RS + slot_of<RB, NW>(KB)];
        // row values of this wave's rows, column values of every block
        RowT<RB, NW, W> rs;
        double cv[RB];
        {
            const double2* p = reinterpret_cast<const double2*>(cp + c.lr * RS + W * H);
#pragma unroll
            for (int m = M0 & ~1; m < O::na; m += 2) {
                const double2 t = p[m / 2];
                if (m >= M0) rs[m] = t.x;
                if (m + 1 < O::na) rs[m + 1 < O::na ? m + 1 : 0] = t.y;
            }
        }
#pragma unroll
        for (int h = 0; h < NW; ++h) {
            const double2* p = reinterpret_cast<const double2*>(cp + c.lc * RS + h * H);
#pragma unroll
            for (int s = 0; s < H; s += 2) {
                const int b0 = h + NW * s, b1 = h + NW * (s + 1);
                const bool u0 = b0 >= KB && b0 < RB, u1 = b1 >= KB && b1 < RB;
                if (u0 || u1) {
                    const double2 t = p[s / 2];
                    if (u0) cv[u0 ? b0 : 0] = t.x;
                    if (u1) cv[u1 ? b1 : 0] = t.y;
                }
            }
        }
        fail = (fail == 0 && !(d > 0.0)) ? k + 1 : fail;
        const double rk = fast_rcp(d);
        if (W == 0 && c.lane == 0) {
            const double y = fast_rsqrt(d);
            c.rdg[k] = y;
            c.dgl[k] = d * y;
        }
#pragma unroll
        for (int m = M0; m < O::na; ++m) rs[m] *= rk;
        // Rows need no mask: an element (i, j) with i <= k < j lies in the
        // unused upper part of a diagonal block, and rows beyond the matrix
        // are never stored.  Finished columns j <= k must not change.
        if (c.lc + 8 * KB <= k) cv[KB] = 0.0;
        fast_update_col<RB, NW, W, KB>(acc, rs, cv);
        if ((k & 7) != 7) {
            fast_publish<RB, NW, W, KB>(acc, c, k + 1, cn);
            if constexpr (KB + 1 < RB) fast_update_col<RB, NW, W, KB + 1>(acc, rs, cv);
        } else {
            if constexpr (KB + 1 < RB) {
                fast_update_col<RB, NW, W, KB + 1>(acc, rs, cv);
                fast_publish<RB, NW, W, KB + 1>(acc, c, k + 1, cn);
            }
        }
        fast_update<RB, NW, W, KB + 2>(acc, rs, cv);
        __syncthreads();
    }
    if constexpr (KB + 1 < RB) {
        if (8 * KB + 8 < c.n) fast_block_steps<RB, NW, W, KB + 1>(acc, c, fail);
    }
}

// Back substitution L^T x = y, columns j = min(8*KB+7, n-1) .. 8*KB
// (descending), then block KB-1: x_j = y_j / L[j,j], then y_r -= L[j,r] x_j
// for r < j by the 8 lanes that hold row j.
template <int RB, int NW, int W, int KB>
__device__ __forceinline__ void fast_back_steps(const AccT<RB, NW, W>& acc, const FastCtx& c,
                                                double* yv, double* xv)
{
    using O = Own<RB, NW, W>;
    const int jhi = (8 * KB + 7 < c.n - 1) ? 8 * KB + 7 : c.n - 1;
    for (int j = jhi; j >= 8 * KB; --j) {
        const double xj = yv[j] * c.rdg[j];
        if (c.tid == 0) xv[j] = xj;
        if constexpr (KB % NW == W) {
            constexpr int m = KB / NW;
            if (c.lr == (j & 7)) {
#pragma unroll
                for (int bb = 0; bb <= KB; ++bb) {
                    const int r = c.lc + 8 * bb;
                    if (r < j) yv[r] -= acc[O::idx(m, bb)] * xj;
                }
            }
        }
        __syncthreads();
    }
    if constexpr (KB > 0) fast_back_steps<RB, NW, W, KB - 1>(acc, c, yv, xv);
}

template <int RB, int NW, int W, int KB>
__device__ __forceinline__ void fast_back_from(const AccT<RB, NW, W>& acc, const FastCtx& c,
                                               double* yv, double* xv, int nb)
{
    // start at the block that holds row n-1
    if (nb == KB) {
        fast_back_steps<RB, NW, W, KB>(acc, c, yv, xv);
        return;
    }
    if constexpr (KB > 0) fast_back_from<RB, NW, W, KB - 1>(acc, c, yv, xv, nb);
}

// One matrix, wave W's share.
template <int RB, int NW, int W, bool POSV>
__device__ __forceinline__ void fast_matrix(const MfhCholBatch& B, FastCtx& c, int b, double* yv, double* xv)
{
    using O = Own<RB, NW, W>;
    const int lr = c.lr, lc = c.lc;
    const int n = __builtin_amdgcn_readfirstlane(B.n[b]);
    const int lda = __builtin_amdgcn_readfirstlane(B.lda[b]);
    c.n = n;
    GPtr A = (GPtr)B.a[b];

    // Load element (i, j), i >= j: the upper-triangle entry A[j + i*lda].
    // Offsets are clamped into the matrix so every load is unconditional,
    // and the value selected.
    double acc[O::size];
#pragma unroll
    for (int m = 0; m < O::na; ++m) {
        const int i = lr + 8 * O::arow(m);
        const int ic = i < n ? i : n - 1;
#pragma unroll
        for (int bb = 0; bb < RB; ++bb) {
            if (bb > O::arow(m)) continue;
            const int j = lc + 8 * bb;
            const int jc = j < ic ? j : ic;
            const double v = gload(A, jc + ic * lda);
            acc[O::idx(m, bb)] = (i < n && j <= i) ? v : 0.0;
        }
        // one block row of loads at a time bounds the live addresses
        __builtin_amdgcn_sched_barrier(0);
    }
    if (POSV) {
        // Row n of the triangle holds b (lanes with lr == n % 8, block n / 8).
        GPtr X = (GPtr)B.x[b];
        double xb[RB];
#pragma unroll
        for (int bb = 0; bb < RB; ++bb) {
            const int j = lc + 8 * bb;
            xb[bb] = gload(X, j < n ? j : n - 1);
        }
        const bool rown = lr == (n & 7);
#pragma unroll
        for (int m = 0; m < O::na; ++m) {
            if (O::arow(m) != (n >> 3)) continue;
#pragma unroll
            for (int bb = 0; bb < RB; ++bb) {
                if (bb > O::arow(m)) continue;
                const int j = lc + 8 * bb;
                acc[O::idx(m, bb)] = (rown && j < n) ? xb[bb] : acc[O::idx(m, bb)];
            }
        }
    }
    fast_publish<RB, NW, W, 0>(acc, c, 0, c.col);
    __syncthreads();
    int fail = 0;
    fast_block_steps<RB, NW, W, 0>(acc, c, fail);
    fail = __builtin_amdgcn_readfirstlane(fail);
    if (fail) {
        // FAST mode leaves a matrix that is not positive definite unchanged.
        if (c.tid == 0) B.info[b] = fail;
        return;
    }
    // Scale the columns: a_ij / L[j,j] (the diagonal is stored from dgl).
#pragma unroll
    for (int bb = 0; bb < RB; ++bb) {
        if (bb > O::arow(O::na - 1)) continue;
        const double r = c.rdg[lc + 8 * bb];
#pragma unroll
        for (int m = 0; m < O::na; ++m)
            if (bb <= O::arow(m)) acc[O::idx(m, bb <= O::arow(m) ? bb : 0)] *= r;
    }
    // Store L, the diagonal and the zeroed strict upper triangle: in a
    // diagonal block each lane writes its own (i, j); off-diagonal blocks
    // write L at (i, j) and 0 at (j, i).  One exec-masked region per block row.
#pragma unroll
    for (int m = 0; m < O::na; ++m) {
        const int a = O::arow(m);
        const int i = lr + 8 * a;
        if (i < n) {
#pragma unroll
            for (int bb = 0; bb < RB; ++bb) {
                if (bb >= a) continue;
                const int j = lc + 8 * bb;
                gstore(A, i + j * lda, acc[O::idx(m, bb)]);
                gstore(A, j + i * lda, 0.0);
            }
            const int j = lc + 8 * a;
            if (j < n) {
                const double v = acc[O::idx(m, a)];
                gstore(A, i + j * lda, (i > j) ? v : (i == j ? c.dgl[i] : 0.0));
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if (c.tid == 0) B.info[b] = 0;
    if (POSV) {
        // y = row n of the triangle, then the Lower/Trans sweep
        // (Blas.SolveTriangular.cs:224-231 order) with x in xv.
        if (lr == (n & 7)) {
#pragma unroll
            for (int m = 0; m < O::na; ++m) {
                if (O::arow(m) != (n >> 3)) continue;
#pragma unroll
                for (int bb = 0; bb < RB; ++bb) {
                    if (bb > O::arow(m)) continue;
                    const int j = lc + 8 * bb;
                    if (j < n) yv[j] = acc[O::idx(m, bb)];
                }
            }
        }
        __syncthreads();
        fast_back_from<RB, NW, W, RB - 1>(acc, c, yv, xv, (n - 1) >> 3);
        GPtr Xo = (GPtr)B.x[b];
        for (int r = c.tid; r < n; r += 64 * NW) gstore(Xo, r, xv[r]);
    }
}

// Waves per SIMD the register allocation of the one-wavefront kernel must
// leave room for.  Where the compiler's own choice lands just above a
// register-file step (256, 168 or 128 registers), a cap that costs a few
// spilled registers doubles (or raises by a half) the matrices in flight
// per CU: at 8*RB = 56 rows 132 -> 128 registers (2 spill instructions), at
// 72 rows 182 -> 168 (6), at 96 rows 284 -> 256 (9), at 104 rows 310 ->
// 256 (268).  Measured on MI355X (8,000-9,000 SPD solves per bucket): 49..56
// rows 0.188 -> 0.170 ms, 65..72 0.315 -> 0.27, 89..96 0.76 (two
// wavefronts) -> 0.50, 97..104 0.63 -> 0.56.  Caps at 64, 80 and 88 rows
// (86 to 869 spill instructions) made those buckets 1.2x to 2.4x slower.
template <int RB, int NW>
__host__ __device__ constexpr int fast_min_waves()
{
    return NW > 1 ? 1 : RB == 7 ? 4 : RB == 9 ? 3 : (RB == 12 || RB == 13) ? 2 : 1;
}

template <int RB, int NW, bool POSV>
__global__ __launch_bounds__(64 * NW, (fast_min_waves<RB, NW>())) void k_chol_fast(MfhCholBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    constexpr int NR = 8 * RB;
    FastCtx c;
    c.col = lds;
    c.dgl = lds + 16 * row_stride<RB, NW>();
    c.rdg = c.dgl + NR;
    double* yv = c.rdg + NR;
    double* xv = yv + NR;
    c.tid = threadIdx.x;
    c.lane = threadIdx.x & 63;
    c.lr = c.lane & 7;
    c.lc = c.lane >> 3;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    for (int it = blockIdx.x;; it += gridDim.x) {
        // One matrix per workgroup: its id and sizes are uniform, which keeps
        // every step and block test below a scalar branch.
        const int b = __builtin_amdgcn_readfirstlane(mfh_item(q, it));
        if (b < 0) break;
        if (NW == 1 || w == 0)
            fast_matrix<RB, NW, 0, POSV>(B, c, b, yv, xv);
        else
            fast_matrix<RB, NW, (NW > 1 ? 1 : 0), POSV>(B, c, b, yv, xv);
        __syncthreads();
    }
}

template <int RB, int NW, bool POSV>
int launch_fast(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, long count)
{
    const size_t lds = (size_t)fast_lds_doubles<RB, NW>() * 8;
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chol_fast<RB, NW, POSV>, 64 * NW, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > count) grid = count;
    if (grid < 1) grid = 1;
    int st = mfh_timing_begin(h, MFH_KERNEL_CHOL_FAST, 8 * RB);
    if (st) return st;
    hipLaunchKernelGGL((k_chol_fast<RB, NW, POSV>), dim3((unsigned)grid), dim3(64 * NW), lds, h->stream, B, q);
    MFH_HIP_CHECK(hipGetLastError());
    return mfh_timing_end(h);
}

template <int RB, bool POSV>
int launch_fast_nw(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, long count)
{
    const int nw = h->fast_waves[RB] > 0 ? h->fast_waves[RB] : mfh_fast_default_waves(RB);
    if constexpr (RB >= 2) {
        if (nw == 2) return launch_fast<RB, 2, POSV>(h, B, q, count);
    }
    return launch_fast<RB, 1, POSV>(h, B, q, count);
}

template <bool POSV>
int dispatch_fast(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count)
{
    switch (rb) {
    case 1: return launch_fast_nw<1, POSV>(h, B, q, count);
    case 2: return launch_fast_nw<2, POSV>(h, B, q, count);
    case 3: return launch_fast_nw<3, POSV>(h, B, q, count);
    case 4: return launch_fast_nw<4, POSV>(h, B, q, count);
    case 5: return launch_fast_nw<5, POSV>(h, B, q, count);
    case 6: return launch_fast_nw<6, POSV>(h, B, q, count);
    case 7: return launch_fast_nw<7, POSV>(h, B, q, count);
    case 8: return launch_fast_nw<8, POSV>(h, B, q, count);
    case 9: return launch_fast_nw<9, POSV>(h, B, q, count);
    case 10: return launch_fast_nw<10, POSV>(h, B, q, count);
    case 11: return launch_fast_nw<11, POSV>(h, B, q, count);
    case 12: return launch_fast_nw<12, POSV>(h, B, q, count);
    case 13: return launch_fast_nw<13, POSV>(h, B, q, count);
    case 14: return launch_fast_nw<14, POSV>(h, B, q, count);
    case 15: return launch_fast_nw<15, POSV>(h, B, q, count);
    case 16: return launch_fast_nw<16, POSV>(h, B, q, count);
    case 17: return launch_fast_nw<17, POSV>(h, B, q, count);
    default: return mfh_fail(MFH_ERR_UNSUPPORTED, "fast Cholesky: too many rows");
    }
}

}  // namespace
