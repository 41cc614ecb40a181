// mfh_fast_kernels.h -- FAST-mode batched FP64 Cholesky / SPD solve for gfx950.
//
// The kernel templates are included by two translation units that build in
// parallel: mfh_fast.hip (Cholesky alone, and the bucket dispatcher) and
// mfh_fast_posv.hip (the SPD solve).
//
// The exact kernels (mfh_exact.hip) reproduce MatFlat's operation order bit
// for bit (Factorization.Cholesky.cs:114-168), which forces one unfused
// multiply and one add per product and a dot-product schedule.  FAST mode
// gives up bit-identity for throughput: results match MatFlat to rounding
// (backward error of order n*eps), checked against the oracle by residual.
//
// Design: blocked right-looking Cholesky, one workgroup of NW (1, 2 or 4)
// wavefronts per matrix of at most 8*RB rows (RB <= 17).
//  * The lower triangle lives in VGPRs, block-cyclic over an 8 x 8 lane
//    grid: lane (lr, lc) of wave w owns elements (lr + 8a, lc + 8b) of the
//    blocks (a, b), b <= a, of its block rows a = w (mod NW).
//  * Panel p (columns 8p .. 8p+7): the owners write its blocks to an LDS
//    panel buffer; wavefront 0 reads it back one row per lane (two or three
//    rows per lane past 64 and 128 rows) and factors the eight columns with v_readlane
//    broadcasts -- the pivot and the panel rows are read from the lanes
//    that hold them, so a column costs no LDS round trip and no barrier;
//    it writes the scaled L panel back.  Every wave then stores its rows of
//    the L panel to HBM straight from LDS (so the writes of L spread over
//    the factorization) and applies the rank-8 update A22 -= L21 L21^T to
//    the blocks it owns, as eight rank-1 rounds whose LDS operand loads are
//    issued together.  Two barriers per panel; the panel buffer is
//    double-buffered by panel parity.  The register tile is never written
//    at a panel-dependent position: it keeps each panel as it stood before
//    factoring, P_Jp, with L_Jp = P_Jp L_pp^-T.
//  * The SPD solve carries b as an extra row n of the triangle: the panel
//    factorization turns that row into y = L^-1 b.  The back substitution
//    L^T x = y runs by block columns K, descending, on the kept P blocks:
//    L_KK^T x_K = y_K - L_KK^-1 z_K with z_K = sum over J > K of P_JK^T x_J.
//    Every wave sums its blocks' part of z_K, the sums meet in LDS (one
//    barrier per block column) and every wave then solves the two 8 x 8
//    triangular systems with the diagonal block kept in LDS, one lane per
//    row and v_readlane broadcasts; x stays in each wave's own LDS copy.
//  * A pivot that is not positive stops the factorization (info = j + 1):
//    the columns of L of the panels before it have been written, x is left
//    unchanged.
//  * Loads of the upper triangle and stores of L (and of the zeroed strict
//    upper triangle) touch each HBM byte once, eight contiguous doubles per
//    lane group, through 32-bit offsets from a wave-uniform base.
#pragma once
#include <type_traits>

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
    static constexpr int na = (RB - W + NW - 1) / NW > 0 ? (RB - W + NW - 1) / NW : 0;
    __host__ __device__ __forceinline__ static constexpr int arow(int m) { return W + NW * m; }
    // sum of arow(t) + 1 over t < m, in closed form
    __host__ __device__ __forceinline__ static constexpr int base(int m)
    {
        return m * (W + 1) + NW * (m * (m - 1) / 2);
    }
    static constexpr int size = base(na) > 0 ? base(na) : 1;
    __host__ __device__ __forceinline__ static constexpr int idx(int m, int bb) { return base(m) + bb; }
};

template <int RB, int NW, int W>
using AccT = double[Own<RB, NW, W>::size];

// LDS layout of a workgroup, in doubles.  A panel buffer holds 8*RB rows of
// eight doubles at a row stride of PLD = 9: both access patterns -- one row
// per lane, and eight rows each read by eight lanes -- then fall on distinct
// banks, and every address is a per-lane base plus an immediate offset (a
// swizzle computed per element would cost one address register each).
constexpr int PLD = 9;
template <int RB, int NW>
struct FastLds {
    static constexpr int NR = 8 * RB;
    static constexpr int PB = PLD * NR;        // one panel buffer
    static constexpr int NPB = NW > 1 ? 2 : 1;  // double-buffered across wavefronts
    static constexpr int rdg = NPB * PB;       // [NR] 1 / L[j][j]
    static constexpr int yv = rdg + NR;        // [NR] y = L^-1 b (SPD solve)
    static constexpr int xw = yv + NR;         // [NW][NR] each wave's copy of x
    static constexpr int part = xw + NW * NR;  // [2][NW][8] block-column sums
    static constexpr int flag = part + 16 * NW;  // failing row + 1 (as a double)
    static constexpr int total = flag + 2;
};

__device__ __forceinline__ int pidx(int r, int t) { return r * PLD + t; }

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
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

// Calls f(std::integral_constant<int, P>) for the uniform value P = p < N:
// one scalar branch tree, and the body sees the panel index as a constant,
// so every register of the tile it names is fixed.  The bodies only read
// the tile (a body that wrote it would leave the compiler merging whole
// copies of the tile after the switch).
template <int N, typename F>
__device__ __forceinline__ void for_panel(int p, F&& f)
{
    switch (p) {
#define MFH_FP_CASE(Q)                                                            \
    case Q:                                                                       \
        if constexpr ((Q) < N) f(std::integral_constant<int, ((Q) < N ? (Q) : 0)>{}); \
        break;
        MFH_FP_CASE(0) MFH_FP_CASE(1) MFH_FP_CASE(2) MFH_FP_CASE(3) MFH_FP_CASE(4) MFH_FP_CASE(5)
        MFH_FP_CASE(6) MFH_FP_CASE(7) MFH_FP_CASE(8) MFH_FP_CASE(9) MFH_FP_CASE(10) MFH_FP_CASE(11)
        MFH_FP_CASE(12) MFH_FP_CASE(13) MFH_FP_CASE(14) MFH_FP_CASE(15) MFH_FP_CASE(16)
#undef MFH_FP_CASE
    default:
        break;
    }
}

// Writes block column P of the owned rows at or below block row P into the
// panel buffer (entry (8a + lr, lc)).
template <int RB, int NW, int W, int P>
__device__ __forceinline__ void publish_p(const AccT<RB, NW, W>& acc, double* pr)
{
    using O = Own<RB, NW, W>;
#pragma unroll
    for (int m = 0; m < O::na; ++m)
        if (O::arow(m) >= P) pr[PLD * 8 * O::arow(m)] = acc[O::idx(m, O::arow(m) >= P ? P : 0)];
}

// A22 -= L21 L21^T on the owned blocks (m, bb), p < bb <= arow(m): eight
// rank-1 rounds; a round's operands (the panel value of each owned row and
// of each block column's row) are loaded from LDS together before its FMAs,
// which run block column by block column under a uniform test.  The tile
// holds -A and the panel buffer -L, so every update is one FMA with no
// negation.
template <int RB, int NW, int W>
__device__ __forceinline__ void trailing(AccT<RB, NW, W>& acc, const double* pr, const double* pc, int p)
{
    using O = Own<RB, NW, W>;
    if constexpr (O::na > 0 && RB > 1) {
        // a loop, not unrolled: the compiler would otherwise hoist the
        // operand loads of every round together
#pragma clang loop unroll(disable)
        for (int t = 0; t < 8; ++t) {
            double rv[O::na];
            double cv[RB];
#pragma unroll
            for (int m = 0; m < O::na; ++m) rv[m] = pr[PLD * 8 * O::arow(m) + t];
#pragma unroll
            for (int bb = 1; bb < RB; ++bb) cv[bb] = pc[PLD * 8 * bb + t];
#pragma unroll
            for (int bb = RB - 1; bb >= 1; --bb) {
                if (bb > p) {
#pragma unroll
                    for (int m = 0; m < O::na; ++m) {
                        if (O::arow(m) >= bb) {
                            double& a = acc[O::idx(m, O::arow(m) >= bb ? bb : 0)];
                            a = __builtin_fma(rv[m], cv[bb], a);
                        }
                    }
                }
            }
        }
    }
}

// Factors panel p in wavefront 0: rows 8p + 64k + lane for k < R (the
// rows of the panel: R = 1, 2 or 3 per lane), columns 8p .. 8p + 7.
// Unscaled right-looking steps with the pivot and the row values broadcast
// by v_readlane; each column is scaled by 1/sqrt(d) as it finishes.  The
// buffer holds the panel negated (-P) and receives -L: the negated rows back
// (rows 8p .. 8p + 7, the diagonal block -L_pp, stay in the buffer to the
// end: later panels write only rows below them), and, for the SPD solve,
// 1/L[j][j] in rdg and the negated row of b (row n: -y, y = L^-1 b) in yv.
// Columns past n (nc <= t < 8) are run with a unit pivot: they hold zeros
// and finite garbage only, and their rdg and yv entries are written as 0.
// Returns the first failing row + 1 (or the fail passed in).
template <int RB, int R, bool POSV>
__device__ __forceinline__ int panel_factor(double* pb, double* rdg, double* yv, int p, int n, int lane, int fail)
{
    constexpr int NR = 8 * RB;
    const int nc = n - 8 * p < 8 ? n - 8 * p : 8;
    double v[R][8];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int r = 8 * p + 64 * k + lane;
        const double* src = pb + PLD * (r < NR ? r : 0);
#pragma unroll
        for (int t = 0; t < 8; ++t) v[k][t] = src[t];
    }
    double ymine = 0.0;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        // the buffer holds -a: d = a_tt, s[u] = -a_ut; each row keeps -a
        // and ends with -L
        double d = -readlane_d(v[0][t], t);
        const bool live = t < nc;
        fail = (fail == 0 && live && !(d > 0.0)) ? 8 * p + t + 1 : fail;
        d = live ? d : 1.0;
        double s[8];
#pragma unroll
        for (int u = t + 1; u < 8; ++u) s[u] = readlane_d(v[0][t], u);
        const double y = fast_rsqrt(d);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const double l = v[k][t] * y;  // -L[r][8p + t]
            const double w = l * y;
            v[k][t] = l;
#pragma unroll
            for (int u = t + 1; u < 8; ++u) v[k][u] = __builtin_fma(w, s[u], v[k][u]);
        }
        ymine = (lane == t) ? (live ? y : 0.0) : ymine;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int r = 8 * p + 64 * k + lane;
        if (r < NR) {
            double* dst = pb + PLD * r;
#pragma unroll
            for (int t = 0; t < 8; ++t) dst[t] = v[k][t];
        }
        if constexpr (POSV) {
            if (r == n) {
#pragma unroll
                for (int t = 0; t < 8; ++t) yv[8 * p + t] = t < nc ? v[k][t] : 0.0;
            }
        }
    }
    if constexpr (POSV) {
        if (lane < 8) rdg[8 * p + lane] = ymine;
    }
    return fail;
}

// Stores the factored panel P of the owned block rows from the panel buffer:
// L at (i, j) and 0 at (j, i) for off-diagonal blocks; in the diagonal block
// L on and below the diagonal and 0 above it.  The row of b (row n) and the
// padding rows are not part of A.  offl = i0 + j*lda and offu = j + i0*lda
// for the lane's row i0 = lr of block row 0 and column j = 8P + lc.
template <int RB, int NW, int W, int P>
__device__ __forceinline__ void store_panel_p(GPtr A, int lda, const double* pr, int offl, int offu,
                                              int nrow, bool jin, int lr, int lc)
{
    using O = Own<RB, NW, W>;
#pragma unroll
    for (int m = 0; m < O::na; ++m) {
        const int a = O::arow(m);
        if (a > P) {
            if (jin && 8 * a < nrow) {
                gstore(A, offl + 8 * a, -pr[PLD * 8 * a]);
                gstore(A, offu + 8 * a * lda, 0.0);
            }
        } else if (a == P) {
            // the diagonal block (element (lr, lc)); above its diagonal the
            // buffer holds leftovers of the unscaled steps
            if (jin && 8 * a < nrow) gstore(A, offl + 8 * a, lr >= lc ? -pr[PLD * 8 * a] : 0.0);
        }
    }
}

// Sum over the eight lanes of a lane group (lanes 8g .. 8g + 7) through DPP:
// quad swaps, then the mirror within each half row.
__device__ __forceinline__ double group8_sum(double v)
{
    auto dpp = [](double x, int ctrl) {
        const int lo = __double2loint(x), hi = __double2hiint(x);
        int rl, rh;
        switch (ctrl) {
        case 0:
            rl = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
            rh = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
            break;
        case 1:
            rl = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
            rh = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false);
            break;
        default:
            rl = __builtin_amdgcn_update_dpp(0, lo, 0x141, 0xF, 0xF, false);  // row_half_mirror
            rh = __builtin_amdgcn_update_dpp(0, hi, 0x141, 0xF, 0xF, false);
            break;
        }
        return __hiloint2double(rh, rl);
    };
    v += dpp(v, 0);
    v += dpp(v, 1);
    v += dpp(v, 2);
    return v;
}

// Contribution of the owned blocks (m, K), arow(m) > K, to -z_K: lane
// (lr, lc) adds (-P)[8a + lr][8K + lc] x[8a + lr], where P is the tile as it
// stood before panel K was factored.
template <int RB, int NW, int W, int K>
__device__ __forceinline__ double colsum_p(const AccT<RB, NW, W>& acc, const double* xr)
{
    using O = Own<RB, NW, W>;
    double s = 0.0;
#pragma unroll
    for (int m = O::na - 1; m >= 0; --m)
        if (O::arow(m) > K) s = __builtin_fma(acc[O::idx(m, O::arow(m) > K ? K : 0)], xr[8 * O::arow(m)], s);
    return s;
}

// Reads through the constant address space: the size, pointer and queue
// arrays are not written while a kernel runs, so uniform indices give
// scalar loads and the values land in scalar registers.
__device__ __forceinline__ int cload_i(const int* p, int i)
{
    return ((const __attribute__((address_space(4))) int*)p)[i];
}
__device__ __forceinline__ GPtr cload_p(double* const* p, int i)
{
    return (GPtr)((double* const __attribute__((address_space(4)))*)p)[i];
}

// One matrix's order, leading dimension and pointers (x: SPD solve only).
struct FastMeta {
    int n, lda;
    GPtr a, x;
};

template <bool POSV>
__device__ __forceinline__ FastMeta fast_meta(const MfhCholBatch& B, int b)
{
    FastMeta m{0, 0, nullptr, nullptr};
    if (b >= 0) {
        m.n = cload_i(B.n, b);
        m.lda = cload_i(B.lda, b);
        m.a = cload_p(B.a, b);
        if constexpr (POSV) m.x = cload_p(B.x, b);
    }
    return m;
}

// One matrix, wave W's share.
template <int RB, int NW, int W, bool POSV>
__device__ __forceinline__ void fast_matrix(const MfhCholBatch& B, const FastMeta& M, double* lds, int b, int lane0)
{
    using O = Own<RB, NW, W>;
    using LY = FastLds<RB, NW>;
    constexpr int NR = 8 * RB;
    const int lr0 = lane0 & 7, lc0 = lane0 >> 3;
    const int n = M.n;
    const int lda = M.lda;
    const int rows = POSV ? n + 1 : n;
    const int np = (n + 7) >> 3;  // panels
    GPtr A = M.a;
    double* rdg = lds + LY::rdg;
    double* yv = lds + LY::yv;

    // Load element (i, j), i >= j: the upper-triangle entry A[j + i*lda].
    // A row past the matrix reads element j of column 0 (in bounds: j < n)
    // and keeps 0; the diagonal block clamps j to i.
    int lr = lr0, lc = lc0;
    asm volatile("" : "+v"(lr), "+v"(lc));
    AccT<RB, NW, W> acc;
#pragma unroll
    for (int m = 0; m < O::na; ++m) {
        const int i = lr + 8 * O::arow(m);
        const bool rin = i < n;
        const int roff = rin ? i * lda : 0;
#pragma unroll
        for (int bb = 0; bb < RB; ++bb) {
            if (bb > O::arow(m)) continue;
            const int j = lc + 8 * bb;
            if (bb < O::arow(m)) {
                const double v = gload(A, j + roff);
                acc[O::idx(m, bb)] = rin ? -v : 0.0;
            } else {
                const int jc = j < i ? j : i;
                const double v = gload(A, (rin ? jc : 0) + roff);
                acc[O::idx(m, bb)] = (rin && j <= i) ? -v : 0.0;
            }
        }
        // one block row of addresses live at a time
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (POSV) {
        // Row n of the triangle holds b (lanes with lr == n % 8, block n / 8).
        GPtr X = M.x;
        const bool rown = lr == (n & 7);
#pragma unroll
        for (int m = 0; m < O::na; ++m) {
            if (O::arow(m) != (n >> 3)) continue;
#pragma unroll
            for (int bb = 0; bb < RB; ++bb) {
                if (bb > O::arow(m)) continue;
                const int j = lc + 8 * bb;
                const double xb = gload(X, j < n ? j : n - 1);
                acc[O::idx(m, bb)] = (rown && j < n) ? -xb : acc[O::idx(m, bb)];
            }
        }
    }
    // per-lane bases of the panel-buffer reads and of the stores of L
    const int offl0 = lr + lc * lda, offu0 = lc + lr * lda;
    const int nrow = n - lr, ncol = n - lc;

    int fail = 0;
    // panel 0's blocks go to the buffer before the loop (the loop publishes
    // panels 1..: one fewer copy of the switch over the tile per iteration,
    // and a few registers fewer)
    publish_p<RB, NW, W, 0>(acc, lds + PLD * lr + lc);
    // not unrolled: one copy of the panel step keeps the code small
#pragma clang loop unroll(disable)
    for (int p = 0; p < np; ++p) {
        double* pb = lds + (LY::NPB > 1 ? (p & 1) * LY::PB : 0);
        const double* pr = pb + PLD * lr + lc;  // rows 8a + lr, column lc (publish, stores)
        if (p > 0) for_panel<RB>(p, [&](auto P) { publish_p<RB, NW, W, decltype(P)::value>(acc, pb + PLD * lr + lc); });
        __syncthreads();
        if constexpr (W == 0) {
            // the panel is the critical path of the whole matrix: its
            // dependent chain issues ahead of other waves' bulk updates
            __builtin_amdgcn_s_setprio(3);
            constexpr int RMAX = (NR + 63) / 64;
            const int rp = rows - 8 * p;  // rows of the panel
            if (RMAX >= 3 && rp > 128)
                fail = panel_factor<RB, (RMAX >= 3 ? 3 : RMAX), POSV>(pb, rdg, yv, p, n, lane0, fail);
            else if (RMAX >= 2 && rp > 64)
                fail = panel_factor<RB, (RMAX >= 2 ? 2 : 1), POSV>(pb, rdg, yv, p, n, lane0, fail);
            else
                fail = panel_factor<RB, 1, POSV>(pb, rdg, yv, p, n, lane0, fail);
            if (lane0 == 0) lds[LY::flag] = (double)fail;
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
        fail = __builtin_amdgcn_readfirstlane((int)lds[LY::flag]);
        if (fail) break;
        const int offl = offl0 + 8 * p * lda, offu = offu0 + 8 * p;
        const bool jin = 8 * p < ncol;
        for_panel<RB>(p, [&](auto P) {
            store_panel_p<RB, NW, W, decltype(P)::value>(A, lda, pr, offl, offu, nrow, jin, lr, lc);
        });
        trailing<RB, NW, W>(acc, pb + PLD * lr, pb + PLD * lc, p);
    }
    if (fail) {
        // columns of L before the failing panel have been written
        if (W == 0 && lane0 == 0) B.info[b] = fail;
        return;
    }
    if (W == 0 && lane0 == 0) B.info[b] = 0;
    if constexpr (POSV) {
        // Back substitution L^T x = y by block columns K, descending.  The
        // tile still holds each panel as it was before it was factored
        // (P_JK, with L_JK = P_JK L_KK^-T), so
        //   L_KK^T x_K = y_K - L_KK^-1 z_K,   z_K = sum over J > K of P_JK^T x_J.
        double* xw = lds + LY::xw + W * NR;
        // a chain of dependent steps per block column
        __builtin_amdgcn_s_setprio(2);
        for (int i = lane0; i < NR; i += 64) xw[i] = 0.0;
        double* part = lds + LY::part;
        const int t = lane0 & 7;
#pragma clang loop unroll(disable)
        for (int K = np - 1; K >= 0; --K) {
            double s = 0.0;
            for_panel<RB>(K, [&](auto KK) { s = colsum_p<RB, NW, W, decltype(KK)::value>(acc, xw + lr); });
            s = group8_sum(s);
            double* pk = part + (K & 1) * 8 * NW;
            if (lr == 0) pk[8 * W + lc] = s;
            if constexpr (NW > 1) __syncthreads();
            // the 8 x 8 diagonal block, one column (and row) per lane t
            double z = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) z += pk[8 * w + t];
            const double rd = rdg[8 * K + t];
            // -L_KK: rows 8K .. 8K + 7 of panel K's buffer
            const double* dg = lds + (LY::NPB > 1 ? (K & 1) * LY::PB : 0) + PLD * 8 * K;
            double lrow[8], lcol[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                lrow[k] = dg[PLD * t + k];  // -L[t][k]
                lcol[k] = dg[PLD * k + t];  // -L[k][t]
            }
            // The tile, dg and yv hold negated values (-P, -L_KK, -y), so z
            // here is -z and the forward sweep gives -w; then
            // y - w = (-w) - (-y).  Columns past n have rd = 0 and yv = 0,
            // so their x is 0.
            double wmine = 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double v = z * rd;
                wmine = (t == k) ? v : wmine;
                z = __builtin_fma(lrow[k], readlane_d(v, k), z);
            }
            double r = wmine - yv[8 * K + t];
            double xmine = 0.0;
#pragma unroll
            for (int k = 7; k >= 0; --k) {
                const double v = r * rd;
                xmine = (t == k) ? v : xmine;
                r = __builtin_fma(lcol[k], readlane_d(v, k), r);
            }
            if (lane0 < 8) xw[8 * K + lane0] = xmine;
        }
        __builtin_amdgcn_s_setprio(0);
        if constexpr (W == 0) {
            GPtr Xo = M.x;
            for (int r = lane0; r < n; r += 64) gstore(Xo, r, xw[r]);
        }
    }
}

template <int RB, int NW, int W, bool POSV>
__device__ __forceinline__ void fast_wave(const MfhCholBatch& B, const FastMeta& M, double* lds, int b, int lane, int w)
{
    if constexpr (W + 1 < NW) {
        if (w == W) fast_matrix<RB, NW, W, POSV>(B, M, lds, b, lane);
        else fast_wave<RB, NW, W + 1, POSV>(B, M, lds, b, lane, w);
    } else {
        fast_matrix<RB, NW, W, POSV>(B, M, lds, b, lane);
    }
}

// Waves per SIMD the register allocation must leave room for, for the
// one-wavefront kernels whose natural allocation sits just above a step of
// the 512-register file: a bucket runs as many matrices at once as its
// wavefronts per SIMD, and its time is per-matrix latency.  Measured on
// MI355X (each bucket alone, 100,000-solve batch): 41..48 rows 5 instead of
// 4 (0.111 -> 0.102 ms), 57..64 rows 4 instead of 3, 65..72 rows 3 instead of
// 2 (0.274 -> 0.240 ms, no spill at 168), 73..80 rows 3 (0.331 -> 0.296 ms,
// 68 bytes spilled), 97..104 rows 2 instead of 1 (0.484 -> 0.335 ms); 17..24
// rows 8.  Capping 25..32, 49..56 and 81..88 rows measured slower (their
// spills cost more than the extra matrices gain).
template <int RB, int NW>
__host__ __device__ constexpr int fast_min_waves()
{
    if constexpr (NW != 1) return 1;
    return RB == 3 ? 8 : RB == 6 ? 5 : RB == 8 ? 4 : RB == 9 ? 3 : RB == 10 ? 3 : RB == 13 ? 2 : 1;
}

template <int RB, int NW, bool POSV>
__global__ __launch_bounds__(64 * NW, (fast_min_waves<RB, NW>())) void k_chol_fast(MfhCholBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // One matrix per workgroup: its id and sizes are uniform, which keeps
    // every panel and block test below a scalar branch.  The ids and sizes
    // are read a matrix ahead: the scalar loads of the next matrix's id and
    // of the following one's sizes and pointers go out beside this matrix's
    // tile loads, so a matrix starts with its tile loads instead of a chain
    // of dependent reads (queue extent, id, sizes and pointers, tile).
    const int qn = q.items ? cload_i(q.count, 0) : q.direct;
    const int q0 = q.items ? cload_i(q.start, 0) : 0;
    const int G = gridDim.x;
    auto item = [&](int idx) -> int { return idx < qn ? (q.items ? cload_i(q.items, q0 + idx) : idx) : -1; };
    int it = blockIdx.x;
    int b = item(it);
    FastMeta m = fast_meta<POSV>(B, b);
    int bn = item(it + G);
    while (b >= 0) {
        const FastMeta mn = fast_meta<POSV>(B, bn);
        const int bnn = item(it + 2 * G);
        fast_wave<RB, NW, 0, POSV>(B, m, lds, b, lane, w);
        __syncthreads();
        b = bn;
        m = mn;
        bn = bnn;
        it += G;
    }
}

template <int RB, int NW, bool POSV>
int launch_fast(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, long count)
{
    const size_t lds = (size_t)FastLds<RB, NW>::total * 8;
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
    if constexpr (RB >= 4) {
        if (nw == 4) return launch_fast<RB, 4, POSV>(h, B, q, count);
    }
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
