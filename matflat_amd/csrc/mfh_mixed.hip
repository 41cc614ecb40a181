// mfh_mixed.hip -- batched LU solve for n <= 512: the mixed-precision solver
// (FP32 factorization, BF16-split MFMA trailing updates for n >= 256, FP64
// iterative refinement) and the plain FP64 blocked solver it is compared
// with.  Neither has a single MatFlat counterpart: they compose
// Factorization.Lu (Factorization.Lu.cs:122-161) with the LU solve of
// Factorization.LuInverse.cs:169-192, and are judged by backward error, not
// bit-identity.
//
// One workgroup of 512 threads (8 wavefronts) per matrix, persistent over the
// batch through an atomic work counter:
//  * A (FP64, column-major) is copied into a row-major workspace W in the
//    factorization precision T, through an LDS-tiled transpose;
//  * right-looking blocked LU with partial pivoting, panel width NB (32 for
//    FP32, 16 for FP64): the panel lives in registers (one row per thread),
//    each column takes one block-wide argmax and one row exchange through
//    LDS; the pivots are applied to the other columns by row swaps (W is
//    row-major, so a swap moves two contiguous rows), and U12 = L11^-1 A12
//    by column-parallel forward substitution on the swapped values in
//    registers, kept transposed in LDS; the trailing update A22 -= L21 U12 runs on MFMA:
//      - FP64: v_mfma_f64_16x16x4f64, 16x16 output tiles, K = 16 in 4 steps;
//      - FP32, n >= 256: v_mfma_f32_16x16x32_bf16 on a three-term BF16
//        split of both operands (x = x1 + x2 + x3), six products summed
//        smallest first, FP32 accumulation: FP32-level accuracy from BF16
//        matrix cores;
//      - FP32, n < 256: FP32 FMA on the vector units;
//  * the solve applies the row permutation, then blocked forward / backward
//    substitution (64-row blocks: a block-parallel GEMV, then one wavefront
//    solving the diagonal block with register-resident rows);
//  * mixed precision: x0 = (LU)^-1 P b, then while the normwise backward
//    error eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) exceeds
//    tol_per_n * n and fewer than max_iter corrections were made,
//    x += (LU)^-1 P (b - A x) -- residual, update and test in FP64, reading
//    the original FP64 A.
#include "mfh_device.h"
#include "mfh_internal.h"
#include "mfh_order.h"

// Neither solver is bit-identical to MatFlat: multiply-adds may fuse.
#pragma clang fp contract(fast)

namespace {

constexpr int BT = 512;  // threads per workgroup
// Workspace header (ints): [0] factorization counter, [1, 103) active
// counts of up to 101 steps, [128, 229) the steps' work counters, then the
// order histogram and cursors.
constexpr int kGesvStepOff = 128, kGesvHistOff = 256, kGesvCursorOff = 384, kGesvHeaderInts = 512;
constexpr int NWAVE = BT / 64;

typedef float float4_t __attribute__((ext_vector_type(4)));
typedef double double4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

template <typename T>
struct Prec;
// NB: panel width; BS: triangular-solve block (rows held in registers by
// one wavefront during the diagonal-block solve).
template <>
struct Prec<float> {
    static constexpr int NB = 32;
    static constexpr int BS = 64;
};
template <>
struct Prec<double> {
    static constexpr int NB = 16;
    static constexpr int BS = 32;
};

struct GesvArgs {
    const int* n;
    const double* const* a;
    const int* lda;
    const double* const* b;
    double* const* x;
    int* iters;  // mixed only
    double* berr;
    int* info;
    int count;
    int max_n;
    int max_iter;
    double tol_per_n;
    char* ws;          // per-matrix workspaces, wstride bytes apart
    long long wstride;
    int ldw;           // row stride of W (elements)
    int* counter;      // work counter of the factorization (zeroed before launch)
    int* active;       // [max_iter + 2]: matrices still active entering step it (zeroed before launch)
    int* stepc;        // [max_iter + 1]: work counter of each solve step (zeroed before launch)
    int* order;        // [count]: matrix ids by decreasing order n (8-row buckets)
    // state kept between the factorization and the refinement launches
    int* perm;         // [count][ldw] position -> original row of A
    double* resid;     // [count][ldw] last residual b - A x
    double* norms;     // [count][2] ||A||_inf, ||b||_inf
    int* zero;         // [count] first zero pivot (1-based) or 0
    int* state;        // [count] 0 active, 1 finished, -1 invalid
    int factor_only;   // no right-hand side: factor, then export the factors
    double* const* lu_out;  // factor_only: where the factors go (lda of a), and ...
    int* const* piv_out;    // ... the row order (position -> original row)
};

// LDS layout (bytes), for row stride L = ldw:
//   ut    [L][NB] T       U12 transposed (k contiguous); also the copy tile
//   l11   [NB][NB] T      L11 transposed (column j contiguous)
//   cand  [2][NWAVE][NB] T, ckey/cidx [2][NWAVE]: pivot search
//   dst   [L] int         final row of each panel row
//   xv, rv [L] double     solution, residual / right-hand side
//   yv    [L] T           solve vector
//   perm  [L] int         row permutation (position -> original row)
//   pivs  [NB] int, mapk / mapv [64] int, mapn
//   red   [NWAVE] (double key, int idx), scal [8] double
template <typename T>
struct Lds {
    T* ut;
    T* l11;
    T* cand;     // [2][NWAVE][NB]: each wavefront's pivot candidate row, by column parity
    double* ckey;  // [2][NWAVE] candidate keys
    int* cidx;     // [2][NWAVE] candidate rows
    double* xv;
    double* rv;
    T* yv;
    int* perm;
    int* pivs;
    int* mapk;  // rows touched by a panel's swaps (at most 64) ...
    int* mapv;  // ... and the row whose content each receives
    int* mapn;  // number of such rows
    int* dst;   // [ldw]: final row of panel row k0 + i
    T* diag;    // [BS][BS + 1]: diagonal block of the triangular solves
    double* rkey;
    int* ridx;
    double* scal;
    int* flag;
};

// The ut region also holds the 128 x 33 staging tile of the input copy.
__host__ __device__ inline size_t ut_elems(int ldw, int nb)
{
    const size_t e = (size_t)ldw * nb;
    return e > 128 * 33 ? e : 128 * 33;
}

template <typename T>
__host__ __device__ size_t gesv_lds_bytes(int ldw)
{
    constexpr int NB = Prec<T>::NB;
    size_t s = 0;
    s += ut_elems(ldw, NB) * sizeof(T);
    s += (size_t)NB * NB * sizeof(T);
    s += 2 * (size_t)NWAVE * NB * sizeof(T);
    s = (s + 15) & ~(size_t)15;
    s += 2 * NWAVE * sizeof(double) + 2 * NWAVE * sizeof(int);
    s = (s + 15) & ~(size_t)15;
    s += 2 * (size_t)ldw * sizeof(double);
    s += (size_t)ldw * sizeof(T);
    s = (s + 15) & ~(size_t)15;
    s += (size_t)ldw * sizeof(int);
    s += (size_t)NB * sizeof(int);
    s += 2 * 64 * sizeof(int) + 16;
    s += (size_t)ldw * sizeof(int);
    s = (s + 15) & ~(size_t)15;
    s += NWAVE * sizeof(double) + NWAVE * sizeof(int) + 8 * sizeof(double) + 16;
    return s;
}

template <typename T>
__device__ __forceinline__ Lds<T> carve(char* base, int ldw)
{
    constexpr int NB = Prec<T>::NB;
    Lds<T> L;
    size_t s = 0;
    L.ut = (T*)(base + s);
    s += ut_elems(ldw, NB) * sizeof(T);
    L.l11 = (T*)(base + s);
    s += (size_t)NB * NB * sizeof(T);
    L.cand = (T*)(base + s);
    s += 2 * (size_t)NWAVE * NB * sizeof(T);
    s = (s + 15) & ~(size_t)15;
    L.ckey = (double*)(base + s);
    s += 2 * NWAVE * sizeof(double);
    L.cidx = (int*)(base + s);
    s += 2 * NWAVE * sizeof(int);
    s = (s + 15) & ~(size_t)15;
    L.xv = (double*)(base + s);
    s += (size_t)ldw * sizeof(double);
    L.rv = (double*)(base + s);
    s += (size_t)ldw * sizeof(double);
    L.yv = (T*)(base + s);
    s += (size_t)ldw * sizeof(T);
    s = (s + 15) & ~(size_t)15;
    L.perm = (int*)(base + s);
    s += (size_t)ldw * sizeof(int);
    L.pivs = (int*)(base + s);
    s += NB * sizeof(int);
    L.mapk = (int*)(base + s);
    s += 64 * sizeof(int);
    L.mapv = (int*)(base + s);
    s += 64 * sizeof(int);
    L.mapn = (int*)(base + s);
    s += 16;
    L.dst = (int*)(base + s);
    s += (size_t)ldw * sizeof(int);
    s = (s + 15) & ~(size_t)15;
    L.rkey = (double*)(base + s);
    s += NWAVE * sizeof(double);
    L.scal = (double*)(base + s);
    s += 8 * sizeof(double);
    L.ridx = (int*)(base + s);
    s += NWAVE * sizeof(int);
    L.flag = (int*)(base + s);
    return L;
}

// Pointers read from the caller's pointer arrays are generic; reading
// through them emits flat loads, which count against the LDS counter as
// well, so every LDS wait would also wait for them.  Data of A, b, x and the
// workspace is in global memory: these casts let the loads be global loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem_w(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// A value every lane already holds, made provably uniform for the compiler:
// loops and branches that contain barriers must depend only on such values
// (a branch the compiler believes divergent may be restructured so that the
// lanes of one wavefront reach the barriers in different orders).
__device__ __forceinline__ double uniform(double v)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Maximum that propagates NaN: a NaN entry makes the norm NaN (and the
// backward error NaN, which is then reported as a failure) instead of being
// dropped by the comparison.
__device__ __forceinline__ double nan_max(double a, double m) { return (a > m || a != a) ? a : m; }

// Block-wide max of a non-negative double, NaN propagating (all threads get
// the result).
template <typename T>
__device__ __forceinline__ double block_max(const Lds<T>& L, double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = nan_max(__shfl_xor(v, off, 64), v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) L.rkey[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = L.rkey[0];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) m = nan_max(L.rkey[w], m);
    return uniform(m);
}

// ------------------------------------------------------------------- copy
// W[r][c] = (T) A[r + c*lda] through 128 x 32 LDS tiles (both sides
// coalesced, eight independent loads per thread and tile), with the row sums
// of |A| on the way.  Returns ||A||_inf and sets amax = max |a_ij| (both
// wave-uniform, NaN propagating).
template <typename T>
__device__ __forceinline__ double copy_in(const double* A, int lda, T* W, int ldw, int n, const Lds<T>& L,
                                          double& amax)
{
    double emax = 0.0;
    T* tile = L.ut;
    const int t = threadIdx.x;
    const int rr = t & 127, cq = t >> 7;  // read: row, column quarter
    const int cc = t & 31, rq = t >> 5;   // write: column, row sixteenth
    double best = 0.0;
    for (int r0 = 0; r0 < n; r0 += 128) {
        const int r = r0 + rr;
        double rowsum = 0.0;
        // the next 32 columns' loads go out before this tile is written
        // through LDS, so 16 loads per thread are in flight
        auto load8 = [&](double (&v)[8], int c0) {
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const int c = c0 + cq + 4 * h;
                v[h] = (r < n && c < n) ? gmem(A)[r + (size_t)c * lda] : 0.0;
            }
        };
        double vn[8];
        load8(vn, 0);
        for (int c0 = 0; c0 < n; c0 += 32) {
            double v[8];
#pragma unroll
            for (int h = 0; h < 8; ++h) v[h] = vn[h];
            if (c0 + 32 < n) load8(vn, c0 + 32);
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                rowsum += __builtin_fabs(v[h]);
                emax = nan_max(__builtin_fabs(v[h]), emax);
                tile[rr * 33 + cq + 4 * h] = (T)v[h];
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const int rl = rq + 16 * h;
                const int r2 = r0 + rl, c2 = c0 + cc;
                if (r2 < n && c2 < n) W[(size_t)r2 * ldw + c2] = tile[rl * 33 + cc];
            }
            __syncthreads();
        }
        // four partial sums per row (one per column quarter)
        double* part = (double*)tile;
        part[cq * 128 + rr] = rowsum;
        __syncthreads();
        if (cq == 0 && r < n) {
            const double sum = part[rr] + part[128 + rr] + part[256 + rr] + part[384 + rr];
            best = nan_max(sum, best);
        }
        __syncthreads();
    }
    amax = block_max(L, emax);
    return block_max(L, best);
}

// --------------------------------------------------------------- panel
// Pivot search key of row r for a column value v: the row order is partial
// pivoting's (largest |v| in single precision, lowest row on ties, NaN below
// every number, non-candidates below NaN), packed so that one unsigned 64-bit
// maximum finds it.
// Pivot key of a candidate row in wavefront w: |v| rounded to FP32, as
// ordered bits + 16, with the low three bits replaced by 7 - w (NaN -> 8, so
// it is chosen only when nothing else is left; not a candidate -> 0).  One
// 32-bit maximum over the keys of all wavefronts then picks the largest
// |v| to 2^-20 relative, the lowest wavefront on ties, and names it.
template <typename T>
__device__ __forceinline__ unsigned pivot_key(bool cand, T v, int w)
{
    const float a = __builtin_fabsf((float)v);
    const unsigned m = (a != a) ? 8u : (__float_as_uint(a) + 16u);
    return cand ? ((m & ~7u) | (unsigned)(7 - w)) : 0u;
}

// 1/d to working precision from the hardware reciprocal and Newton steps
// (one for FP32, two for FP64): a few dependent instructions on the panel's
// critical path instead of an IEEE division.
__device__ __forceinline__ float recip(float d)
{
    float r = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(__builtin_fmaf(-d, r, 1.0f), r, r);
}
__device__ __forceinline__ double recip(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    return __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
}

// (lanes the row mask leaves out read 0, the identity of the unsigned
// maximum, so the move and the maximum can fuse into one DPP instruction)
#define MFH_DPP_MAX_U32(v, CTRL, RMASK)                                                 \
    do {                                                                              \
        const unsigned o_ = __builtin_amdgcn_update_dpp(0u, (v), CTRL, RMASK, 0xF, false); \
        (v) = o_ > (v) ? o_ : (v);                                                    \
    } while (0)

// Wave-wide maximum through DPP row shifts and row broadcasts (no LDS
// round trips); every lane receives it.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    MFH_DPP_MAX_U32(v, 0x111, 0xF);  // row_shr:1
    MFH_DPP_MAX_U32(v, 0x112, 0xF);  // row_shr:2
    MFH_DPP_MAX_U32(v, 0x114, 0xF);  // row_shr:4
    MFH_DPP_MAX_U32(v, 0x118, 0xF);  // row_shr:8  -> lane 15 of each row: row maximum
    MFH_DPP_MAX_U32(v, 0x142, 0xA);  // row_bcast:15 into rows 1 and 3
    MFH_DPP_MAX_U32(v, 0x143, 0xC);  // row_bcast:31 into rows 2 and 3 -> lane 63: maximum
    return __builtin_amdgcn_readlane(v, 63);
}

// Column J of the panel (J a template parameter, so every register index is
// static; a runtime column loop would index the row registers dynamically).
// Rows stay in their threads: a row chosen as a pivot stops taking part
// ("live" goes false) instead of being exchanged, and the final row order is
// applied once, when the panel is written back.  One barrier per column:
// before it each wavefront's argmax winner publishes its entries J.. of the
// row (double-buffered by column parity); after it every live row eliminates
// with the pivot row.  pivs[J] records the original row chosen at column J.
template <typename T, int J>
__device__ __forceinline__ void panel_cols(T (&P)[Prec<T>::NB], int r, bool& live, int n, int k0, int nb,
                                           const Lds<T>& L, int& zero)
{
    constexpr int NB = Prec<T>::NB;
    if constexpr (J < NB) {
        if (J < nb) {
            const int t = threadIdx.x, w = t >> 6;
            T* cand = L.cand + (J & 1) * NWAVE * NB;
            unsigned* ckey = (unsigned*)(L.ckey + (J & 1) * NWAVE);
            int* cidx = L.cidx + (J & 1) * NWAVE;
            // wavefront 0 always has rows (k0 < n)
            const int nwl = (n - k0 + 63) >> 6;  // wavefronts holding rows k0 .. n-1
            // Every wavefront runs the column, also those without rows in
            // k0 .. n-1 (their rows are not live: key 0, masked below, and
            // l = 0): a branch around it is divergent to the compiler, which
            // then copies the panel row in and out of it and serializes the
            // pivot-row loads behind the copies.
            {
                // wave argmax: the largest 32-bit key, then the lowest lane
                // that holds it (rows are in lane order, so the lowest row
                // wins ties)
                const unsigned k32 = pivot_key(live, P[J], w);
                const unsigned m = wave_max_u32(k32);
                const int win = __builtin_ctzll(__ballot(k32 == m));
                if ((t & 63) == win) {
                    // 16-byte stores from the aligned entry at or below J
                    // (the entries below J are never read): NB / VW - J / VW
                    // stores with immediate offsets
                    typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
                    constexpr int VW = 16 / sizeof(T);
                    vec_t* cv = (vec_t*)(cand + w * NB);
#pragma unroll
                    for (int q = J / VW; q < NB / VW; ++q) {
                        vec_t v;
#pragma unroll
                        for (int e = 0; e < VW; ++e) v[e] = P[q * VW + e];
                        cv[q] = v;
                    }
                    ckey[w] = m;
                    cidx[w] = r;
                }
            }
            __syncthreads();
            {
                // every wavefront's key in one LDS round trip (a loop over
                // the nwl live slots would wait for each load in turn); the
                // slots of wavefronts without rows are masked
                unsigned best = 0u;
#pragma unroll
                for (int ww = 0; ww < NWAVE; ++ww) {
                    const unsigned kw = ckey[ww];
                    best = (ww < nwl && kw > best) ? kw : best;
                }
                best = __builtin_amdgcn_readfirstlane(best);
                const int wb = 7 - (int)(best & 7u);  // the winning wavefront
                const T* prow = cand + wb * NB;
                // a live row always wins (n - k0 - J >= 1 rows are live); the
                // clamp only keeps the index in range whatever the keys hold
                int p = cidx[wb];
                // the pivot row's entries, all loaded before any is used:
                // left to itself the scheduler pairs each load with its FMA
                // and waits for every pair in turn
                T pr[NB];
#pragma unroll
                for (int c = J; c < NB; ++c) pr[c] = prow[c];
                __builtin_amdgcn_sched_barrier(0);
                p = p < k0 ? k0 : (p >= n ? n - 1 : p);
                const T d = pr[J];
                if (t == 0) L.pivs[J] = p;
                if (d == (T)0 && zero == 0) zero = k0 + J + 1;
                if (r == p) live = false;
                // Branch-free: a row that is not live eliminates with l = 0 and
                // keeps its values (a divergent branch here makes the compiler
                // copy the whole row in and out of the branch).  The pivot row's
                // entries are finite unless A holds inf or NaN, when the factors
                // carry NaN whichever way they are computed.
                const T q = (d != (T)0) ? P[J] * recip(d) : P[J];
                const T l = live ? q : (T)0;
                P[J] = live ? q : P[J];
#pragma unroll
                for (int c = J + 1; c < NB; ++c) P[c] -= l * pr[c];
            }
            panel_cols<T, J + 1>(P, r, live, n, k0, nb, L, zero);
        }
    }
}

// Final row order of a panel (wavefront 0).  The pivot of column J (original
// row pivs[J]) goes to row k0+J; the rows of k0..k0+nb-1 that were not
// chosen move, in order, to the rows below the block that the pivots came
// from.  This is a permutation of S = {k0..k0+nb-1} u {pivs[J]} (at most
// 2 nb <= 64 rows), stored as a map: row mapk[i] receives the old content of
// row mapv[i].  It also updates perm (position -> original row of A) and
// dst (final row of each row k0 + i; dst must hold the identity on entry).
template <typename T>
__device__ __forceinline__ void compose_rows(int k0, int nb, const Lds<T>& L)
{
    const int lane = threadIdx.x;
    const int q = lane < nb ? L.pivs[lane] : -1;
    const bool below = lane < nb && q >= k0 + nb;  // pivot taken from below the block
    bool chosen = false;                           // row k0 + lane was some column's pivot
    for (int j = 0; j < nb; ++j) chosen |= __builtin_amdgcn_readlane(q, j) == k0 + lane;
    const bool displaced = lane < nb && !chosen;
    const unsigned long long lower = (1ull << lane) - 1;
    const int rv = __popcll(__ballot(below) & lower), rd = __popcll(__ballot(displaced) & lower);
    const int nd = __popcll(__ballot(displaced));
    if (lane < nb) {
        L.mapk[lane] = k0 + lane;
        L.mapv[lane] = q;
    }
    if (below) L.mapk[nb + rv] = q;
    if (displaced) L.mapv[nb + rd] = k0 + lane;
    const int cnt = nb + nd;
    if (lane == 0) *L.mapn = cnt;
    __builtin_amdgcn_wave_barrier();
    int mk = 0, mv = 0, old = 0;
    if (lane < cnt) {
        mk = L.mapk[lane];
        mv = L.mapv[lane];
        old = L.perm[mv];
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < cnt) {
        L.perm[mk] = old;
        L.dst[mv - k0] = mk;
    }
}

// Factor columns k0 .. k0+nb-1 of rows k0 .. n-1 (one row per thread, in
// registers), compose the panel's row order (mapk/mapv, perm), write the
// panel rows back to their final rows, and leave the unit-lower L11 in l11
// (by columns: L11[i][j] at l11[j NB + i]).
// Returns the first zero pivot (1-based) or 0.
template <typename T>
__device__ __forceinline__ int panel(T* W, int ldw, int n, int k0, int nb, const Lds<T>& L)
{
    constexpr int NB = Prec<T>::NB;
    const int t = threadIdx.x;
    const int r = k0 + t;
    const bool mine = r < n;
    T P[NB];
    {
        // clamped addresses, selected values: branch-free loads.  A full
        // panel is read as 16-byte vectors (rows are 16-byte aligned: ldw and
        // k0 are multiples of 8): every load instruction touches one cache
        // line per lane, so a scalar load per entry would cost the address
        // unit NB passes over the same 64 lines instead of NB / VW.
        const int rc = mine ? r : n - 1;
        if (nb == NB) {
            typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
            constexpr int VW = 16 / sizeof(T);
            const vec_t* src = (const vec_t*)(W + (size_t)rc * ldw + k0);
#pragma unroll
            for (int q = 0; q < NB / VW; ++q) {
                const vec_t v = gmem(src)[q];
#pragma unroll
                for (int e = 0; e < VW; ++e) P[q * VW + e] = mine ? v[e] : (T)0;
            }
        } else {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const T v = W[(size_t)rc * ldw + (k0 + c < n ? k0 + c : n - 1)];
                P[c] = (mine && c < nb) ? v : (T)0;
            }
        }
    }
    int zero = 0;
    bool live = mine;
    panel_cols<T, 0>(P, r, live, n, k0, nb, L, zero);
    if (mine) L.dst[t] = r;
    __syncthreads();
    if (t < 64) compose_rows<T>(k0, nb, L);
    __syncthreads();
    if (mine) {
        const int d = L.dst[t];
        if (nb == NB) {
            typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
            constexpr int VW = 16 / sizeof(T);
            vec_t* dstp = (vec_t*)(W + (size_t)d * ldw + k0);
#pragma unroll
            for (int q = 0; q < NB / VW; ++q) {
                vec_t v;
#pragma unroll
                for (int e = 0; e < VW; ++e) v[e] = P[q * VW + e];
                gmem_w(dstp)[q] = v;
            }
        } else {
            for (int c = 0; c < nb; ++c) {
                T v = P[0];
#pragma unroll
                for (int q = 1; q < NB; ++q) v = (c == q) ? P[q] : v;
                W[(size_t)d * ldw + k0 + c] = v;
            }
        }
        if (d < k0 + nb) {
#pragma unroll
            for (int c = 0; c < NB; ++c) L.l11[c * NB + (d - k0)] = P[c];
        }
    }
    return zero;
}

// Apply the panel's row order (the map composed by compose_rows) to the
// columns outside the panel, and, when a trailing part exists, solve
// U12 = L11^-1 A12 for the columns right of the panel on the way: every
// thread moves its column through registers, all loads first (entries past
// the map's end repeat its last entry, so the 2 NB loads are branch-free and
// go out together); rows k0 .. k0+NB-1 of a column (the first NB entries of
// the map) are A12's column after the swap, so the substitution runs on the
// loaded values before they are stored, and U12 is also written, transposed,
// to ut.  Only full panels have a trailing part (nb == NB there).
template <typename T>
__device__ __forceinline__ void swap_trsm(T* W, int ldw, int n, int k0, int nb, const Lds<T>& L)
{
    constexpr int NB = Prec<T>::NB;
    constexpr int M = 2 * NB;  // compose_rows' map holds at most 2 nb rows
    const int t = threadIdx.x;
    __syncthreads();
    const int cnt = __builtin_amdgcn_readfirstlane(*L.mapn);
    const int c = t;  // n <= BT: one column per thread
    const int c0 = k0 + NB;
    if (c < n && (c < k0 || c >= k0 + nb)) {
        T v[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int e = i < cnt ? i : cnt - 1;
            v[i] = gmem(W)[(unsigned)(L.mapv[e] * ldw + c)];
        }
        // a wavefront holding a column right of the panel (the test is
        // uniform); its lanes left of the panel run the same substitution
        // with every multiplier zero, so their values stay as loaded
        if (k0 + nb < n && wave_id() * 64 + 63 >= c0) {
            const bool right = c >= c0;
            // column-oriented substitution: column j of L11 is contiguous,
            // read as 16-byte vectors from the one holding row j + 1 on
            // (broadcast: every thread reads the same words); the compiler
            // fence (which takes the L11 address) keeps the loads of one
            // column in flight at a time
            typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
            constexpr int VW = 16 / sizeof(T);
#pragma unroll
            for (int j = 0; j < NB - 1; ++j) {
                T lc[NB];
                const vec_t* col = (const vec_t*)(L.l11 + j * NB);
#pragma unroll
                for (int q = (j + 1) / VW; q < NB / VW; ++q) {
                    const vec_t lv = col[q];
#pragma unroll
                    for (int e = 0; e < VW; ++e) lc[q * VW + e] = lv[e];
                }
                const T m = right ? v[j] : (T)0;
#pragma unroll
                for (int i = j + 1; i < NB; ++i) v[i] -= lc[i] * m;
                asm volatile("" ::"v"(L.l11) : "memory");
            }
            if (right) {
#pragma unroll
                for (int i = 0; i < NB; ++i) L.ut[(size_t)(c - c0) * NB + i] = v[i];
            }
        }
#pragma unroll
        for (int i = 0; i < M; ++i)
            if (i < cnt) gmem_w(W)[(unsigned)(L.mapk[i] * ldw + c)] = v[i];
    }
    __syncthreads();
}

// ------------------------------------------------------- trailing updates
// A22 -= L21 U12.  Each wavefront takes a 32 x 32 block (2 x 2 MFMA tiles of
// 16 x 16) per step: the block's C values, L21 rows and U12 columns are all
// loaded up front (independent loads in flight), then the MFMAs run.
struct F64Blk {
    double a[2][4], b[2][4], c[2][2][4];
};

__device__ __forceinline__ void f64_load(F64Blk& B, const double* W, int ldw, int n, int k0, int R0, int row0,
                                         int col0, const Lds<double>& L)
{
    constexpr int NB = Prec<double>::NB;
    const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
    // lane supplies k = 4*kq + s to MFMA step s, for both operands
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int ar = row0 + 16 * h + i;
        const double4_t ta = *(const double4_t*)(W + (size_t)(ar < n ? ar : n - 1) * ldw + k0 + 4 * kq);
        const int bc = col0 + 16 * h + i;
        const double4_t tb = *(const double4_t*)(L.ut + (size_t)((bc < n ? bc : n - 1) - R0) * NB + 4 * kq);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            B.a[h][s] = ar < n ? ta[s] : 0.0;
            B.b[h][s] = bc < n ? tb[s] : 0.0;
        }
    }
    // item r of lane l is row (l/16) + 4r, column l%16
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + 16 * ti + kq + 4 * r, col = col0 + 16 * tj + i;
                B.c[ti][tj][r] = W[(size_t)(row < n ? row : n - 1) * ldw + (col < n ? col : n - 1)];
            }
}

__device__ __forceinline__ void f64_compute_store(const F64Blk& B, double* W, int ldw, int n, int row0, int col0)
{
    const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(B.a[ti][s], B.b[tj][s], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + 16 * ti + kq + 4 * r, col = col0 + 16 * tj + i;
                if (row < n && col < n) W[(size_t)row * ldw + col] = B.c[ti][tj][r] - acc[r];
            }
        }
}

// Software-pipelined like the BF16 variant below.
__device__ __forceinline__ void trailing_f64(double* W, int ldw, int n, int k0, const Lds<double>& L)
{
    constexpr int NB = Prec<double>::NB;  // 16 = 4 MFMA steps of K = 4
    const int R0 = k0 + NB;
    const int m = n - R0;
    const int nt = (m + 31) / 32;
    const int nblk = nt * nt;
    int blk = wave_id();
    if (blk < nblk) {
        F64Blk cur;
        f64_load(cur, W, ldw, n, k0, R0, R0 + (blk / nt) * 32, R0 + (blk % nt) * 32, L);
        for (;;) {
            const int nxt = blk + NWAVE;
            F64Blk nb;
            if (nxt < nblk) f64_load(nb, W, ldw, n, k0, R0, R0 + (nxt / nt) * 32, R0 + (nxt % nt) * 32, L);
            f64_compute_store(cur, W, ldw, n, R0 + (blk / nt) * 32, R0 + (blk % nt) * 32);
            if (nxt >= nblk) break;
            cur = nb;
            blk = nxt;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void split3(float x, __bf16& h1, __bf16& h2, __bf16& h3)
{
    h1 = (__bf16)x;
    const float r1 = x - (float)h1;
    h2 = (__bf16)r1;
    const float r2 = r1 - (float)h2;
    h3 = (__bf16)r2;
}

__device__ __forceinline__ void split8(const float (&x)[8], bf16x8_t& h1, bf16x8_t& h2, bf16x8_t& h3)
{
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        __bf16 a, b, c;
        split3(x[s], a, b, c);
        h1[s] = a;
        h2[s] = b;
        h3[s] = c;
    }
}

// Operands of one 32 x 32 block: L21 rows (a), U12 columns (b) and C.
struct Bf16Blk {
    float a[2][8], b[2][8], c[2][2][4];
};

__device__ __forceinline__ void bf16_load(Bf16Blk& B, const float* W, int ldw, int n, int k0, int R0, int row0,
                                          int col0, const Lds<float>& L)
{
    constexpr int NB = Prec<float>::NB;
    const int lane = threadIdx.x & 63, i = lane & 15, k8 = lane >> 4;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int ar = row0 + 16 * h + i;
        const float4_t* pa = (const float4_t*)(W + (size_t)(ar < n ? ar : n - 1) * ldw + k0 + 8 * k8);
        const float4_t a0 = pa[0], a1 = pa[1];
        const int bc = col0 + 16 * h + i;
        const float4_t* pb = (const float4_t*)(L.ut + (size_t)((bc < n ? bc : n - 1) - R0) * NB + 8 * k8);
        const float4_t b0 = pb[0], b1 = pb[1];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            B.a[h][s] = ar < n ? a0[s] : 0.0f;
            B.a[h][4 + s] = ar < n ? a1[s] : 0.0f;
            B.b[h][s] = bc < n ? b0[s] : 0.0f;
            B.b[h][4 + s] = bc < n ? b1[s] : 0.0f;
        }
    }
    // item r of lane l is row 4*(l/16) + r, column l%16
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + 16 * ti + 4 * k8 + r, col = col0 + 16 * tj + i;
                B.c[ti][tj][r] = W[(size_t)(row < n ? row : n - 1) * ldw + (col < n ? col : n - 1)];
            }
}

__device__ __forceinline__ void bf16_compute_store(const Bf16Blk& B, float* W, int ldw, int n, int row0, int col0)
{
    const int lane = threadIdx.x & 63, i = lane & 15, k8 = lane >> 4;
    bf16x8_t a1[2], a2[2], a3[2], b1[2], b2[2], b3[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        split8(B.a[h], a1[h], a2[h], a3[h]);
        split8(B.b[h], b1[h], b2[h], b3[h]);
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            // six products, smallest first, FP32 accumulation
            float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[ti], b1[tj], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[ti], b2[tj], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[ti], b3[tj], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[ti], b1[tj], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[ti], b2[tj], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[ti], b1[tj], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + 16 * ti + 4 * k8 + r, col = col0 + 16 * tj + i;
                if (row < n && col < n) W[(size_t)row * ldw + col] = B.c[ti][tj][r] - acc[r];
            }
        }
}

// Software-pipelined: the next block's operands are loaded before the
// current block is computed and stored (blocks never overlap).
__device__ __forceinline__ void trailing_bf16x6(float* W, int ldw, int n, int k0, const Lds<float>& L)
{
    constexpr int NB = Prec<float>::NB;  // 32 = K of one MFMA
    const int R0 = k0 + NB;
    const int m = n - R0;
    const int nt = (m + 31) / 32;
    const int nblk = nt * nt;
    int blk = wave_id();
    if (blk < nblk) {
        Bf16Blk cur;
        bf16_load(cur, W, ldw, n, k0, R0, R0 + (blk / nt) * 32, R0 + (blk % nt) * 32, L);
        for (;;) {
            const int nxt = blk + NWAVE;
            Bf16Blk nb;
            if (nxt < nblk) bf16_load(nb, W, ldw, n, k0, R0, R0 + (nxt / nt) * 32, R0 + (nxt % nt) * 32, L);
            bf16_compute_store(cur, W, ldw, n, R0 + (blk / nt) * 32, R0 + (blk % nt) * 32);
            if (nxt >= nblk) break;
            cur = nb;
            blk = nxt;
        }
    }
    __syncthreads();
}

// FP32 FMA on the vector units: each lane computes 4 outputs of a 16 x 16 tile.
__device__ __forceinline__ void trailing_f32_valu(float* W, int ldw, int n, int k0, const Lds<float>& L)
{
    constexpr int NB = Prec<float>::NB;
    const int R0 = k0 + NB;
    const int m = n - R0;
    const int nt = (m + 15) / 16;
    const int lane = threadIdx.x & 63, i = lane & 15, k8 = lane >> 4;
    for (int tile = wave_id(); tile < nt * nt; tile += NWAVE) {
        const int row0 = R0 + (tile / nt) * 16, col0 = R0 + (tile % nt) * 16;
        const int col = col0 + i;
        if (col >= n) continue;
        const float* u = L.ut + (size_t)(col - R0) * NB;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = row0 + 4 * k8 + r;
            if (row >= n) continue;
            const float* l = W + (size_t)row * ldw + k0;
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < NB; ++k) s = __builtin_fmaf(l[k], u[k], s);
            W[(size_t)row * ldw + col] -= s;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------- solve
// The value lane j (a constant) holds, for every lane.
__device__ __forceinline__ float readlane_t(float v, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double readlane_t(double v, int j)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
}

// sum over c in [c0, c1) of row[c] * y[c], split over TPR threads (this one
// is q): thread q takes the VW-entry chunks at c0 + q VW + k TPR VW, loaded
// as 16-byte vectors, GU chunks in flight (c0 is a multiple of TPR VW, so
// every chunk is aligned); a chunk that crosses c1 is summed entry by entry.
// GU = 4 keeps the solve launch within its 80-register budget.
#ifndef MFH_GEMV_U
#define MFH_GEMV_U 4
#endif
template <typename T, int TPR>
__device__ __forceinline__ double gemv_row(const T* row, const T* y, int c0, int c1, int q)
{
    typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int VW = 16 / sizeof(T);
    constexpr int S = TPR * VW;
    constexpr int GU = MFH_GEMV_U;
    double s[GU];
#pragma unroll
    for (int u = 0; u < GU; ++u) s[u] = 0.0;
    int c = c0 + q * VW;
    for (; c + (GU - 1) * S + VW <= c1; c += GU * S) {
        vec_t rv[GU];
#pragma unroll
        for (int u = 0; u < GU; ++u) rv[u] = *(const __attribute__((address_space(1))) vec_t*)(gmem(row) + c + u * S);
        __builtin_amdgcn_sched_barrier(0);  // the loads go out together
#pragma unroll
        for (int u = 0; u < GU; ++u)
#pragma unroll
            for (int e = 0; e < VW; ++e) s[u] += (double)rv[u][e] * (double)y[c + u * S + e];
    }
    for (; c < c1; c += S) {
        if (c + VW <= c1) {
            const vec_t v = *(const __attribute__((address_space(1))) vec_t*)(gmem(row) + c);
#pragma unroll
            for (int e = 0; e < VW; ++e) s[e % GU] += (double)v[e] * (double)y[c + e];
        } else {
            for (int e = 0; c + e < c1; ++e) s[0] += (double)gmem(row)[c + e] * (double)y[c + e];
        }
    }
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < GU; ++u) t += s[u];
    return t;
}

// yv <- U^-1 L^-1 yv with the factors in W (row-major, unit-lower L), in
// blocks of BS rows: a block-parallel GEMV with the finished part (BT/BS
// threads per row), then wavefront 0 solves the diagonal block with its rows
// in registers.
template <typename T>
__device__ __forceinline__ void lu_solve(const T* W, int ldw, int n, const Lds<T>& L)
{
    constexpr int BS = Prec<T>::BS;
    constexpr int TPR = BT / BS;
    constexpr int DL = BS + 1;  // leading dimension of the staged diagonal block
    const int t = threadIdx.x, lane = t & 63;
    const int w = wave_id();
    const int nblk = (n + BS - 1) / BS;
    T* D = L.diag;
    // All threads stage the diagonal block of rows r0 .. r0+rows-1 into LDS
    // (zero outside the triangle the sweep uses), so wavefront 0 solves it
    // from LDS instead of holding it in registers.
    auto stage = [&](int r0, int rows, bool lower) {
        for (int e = t; e < BS * BS; e += BT) {
            const int i = e / BS, j = e - i * BS;
            const bool in = i < rows && j < rows && (lower ? j < i : j >= i);
            const int ri = r0 + (i < rows ? i : rows - 1), cj = r0 + (j < rows ? j : rows - 1);
            const T v = gmem(W)[(size_t)ri * ldw + cj];
            D[i * DL + j] = in ? v : (T)0;
        }
    };
    // forward: unit lower
    for (int blk = 0; blk < nblk; ++blk) {
        const int r0 = blk * BS;
        const int rows = n - r0 < BS ? n - r0 : BS;
        stage(r0, rows, true);
        if (r0 > 0) {
            const int r = r0 + t / TPR, q = t % TPR;
            double s = 0.0;
            if (r < r0 + rows) s = gemv_row<T, TPR>(W + (size_t)r * ldw, L.yv, 0, r0, q);
#pragma unroll
            for (int off = 1; off < TPR; off <<= 1) s += __shfl_xor(s, off, 64);
            if (q == 0 && r < r0 + rows) L.yv[r] = (T)((double)L.yv[r] - s);
        }
        __syncthreads();
        if (w == 0) {
            // j is wave-uniform, so lane j's value reaches the others
            // through v_readlane (lane index in a scalar register) instead
            // of an LDS permute
            const int r = r0 + lane;
            T y = (lane < rows) ? L.yv[r0 + (lane < rows ? lane : 0)] : (T)0;
            const T* drow = D + lane * DL;
#pragma unroll 8
            for (int j = 0; j < BS; ++j) {
                const T yj = readlane_t(y, j);
                y -= drow[j] * yj;
            }
            if (lane < rows) L.yv[r] = y;
        }
        __syncthreads();
    }
    // backward: upper with its diagonal
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int r0 = blk * BS;
        const int rows = n - r0 < BS ? n - r0 : BS;
        const int c1 = r0 + rows;
        stage(r0, rows, false);
        if (c1 < n) {
            const int r = r0 + t / TPR, q = t % TPR;
            double s = 0.0;
            if (r < c1) s = gemv_row<T, TPR>(W + (size_t)r * ldw, L.yv, c1, n, q);
#pragma unroll
            for (int off = 1; off < TPR; off <<= 1) s += __shfl_xor(s, off, 64);
            if (q == 0 && r < c1) L.yv[r] = (T)((double)L.yv[r] - s);
        }
        __syncthreads();
        if (w == 0) {
            // each lane's reciprocal of its diagonal entry first (one division
            // for the wavefront instead of one per column); lane j scales by
            // it when its value is complete
            const int r = r0 + lane;
            T y = (lane < rows) ? L.yv[r0 + (lane < rows ? lane : 0)] : (T)0;
            const T* drow = D + lane * DL;
            const T dd = drow[lane < BS ? lane : 0];
            const T rd = (dd != (T)0) ? (T)1 / dd : (T)1;
#pragma unroll 8
            for (int j = BS - 1; j >= 0; --j) {
                if (lane == j) y *= rd;
                const T yj = readlane_t(y, j);
                if (lane < j) y -= drow[j] * yj;
            }
            if (lane < rows) L.yv[r] = y;
        }
        __syncthreads();
    }
}

// rv <- b - A x (FP64, column-major A); returns ||rv||_inf (all threads).
#ifndef MFH_RESID_U
#define MFH_RESID_U 8
#endif

template <typename T>
__device__ __forceinline__ double residual(const double* A, int lda, const double* b, int n, const Lds<T>& L)
{
    // G groups of threads split the columns (G = 4 for n <= 128, 2 for
    // n <= 256), so every thread has a row; RU independent partial sums keep
    // RU column loads in flight per thread.  The group sums meet
    // in the ut region, which is free once the factorization is done.
    const int G = n <= BT / 4 ? 4 : (n <= BT / 2 ? 2 : 1);
    const int span = BT / G;
    const int t = threadIdx.x;
    const int g = t / span, r = t % span;
    const int cper = (n + G - 1) / G;
    const int c0 = g * cper, c1 = c0 + cper < n ? c0 + cper : n;
    constexpr int RU = MFH_RESID_U;
    double ps[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) ps[u] = 0.0;
    if (r < n) {
        const auto* Ar = gmem(A + r);
        int c = c0;
        for (; c + RU <= c1; c += RU) {
            // all RU loads are issued before the first use: the scheduler
            // would otherwise interleave them with the FMAs and wait for
            // each load in turn
            double av[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u) av[u] = Ar[(size_t)(c + u) * lda];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < RU; ++u) ps[u] = __builtin_fma(-av[u], L.xv[c + u], ps[u]);
        }
        for (; c < c1; ++c) ps[0] = __builtin_fma(-Ar[(size_t)c * lda], L.xv[c], ps[0]);
    }
#pragma unroll
    for (int w = RU / 2; w >= 1; w >>= 1)
#pragma unroll
        for (int u = 0; u < w; ++u) ps[u] += ps[u + w];
    double* part = (double*)L.ut;  // G * n <= BT doubles
    if (r < n) part[g * n + r] = ps[0];
    __syncthreads();
    double m = 0.0;
    for (int rr = t; rr < n; rr += BT) {
        double s = gmem(b)[rr];
        for (int q = 0; q < G; ++q) s += part[q * n + rr];
        L.rv[rr] = s;
        const double a = __builtin_fabs(s);
        m = nan_max(a, m);
    }
    return block_max(L, m);
}

// One matrix (bi), all threads of the workgroup.
// Factorization of matrix bi: norms, the copy into W, and the blocked LU.
// The row order, the norms and the first zero pivot are kept for the
// refinement launches.
template <typename T, bool MIXED>
__device__ __forceinline__ void gesv_factor(const GesvArgs& G, const Lds<T>& L, int bi, int n, int lda)
{
    constexpr int NB = Prec<T>::NB;
    const int ldw = G.ldw;
    const int t = threadIdx.x;
    const double* A = G.a[bi];
    const double* bvec = G.factor_only ? nullptr : G.b[bi];
    T* W = (T*)(G.ws + (long long)bi * G.wstride);

    // ||b||_inf; ||A||_inf comes with the copy
    double bn = 0.0;
    for (int r = t; r < (G.factor_only ? 0 : n); r += BT) {
        const double v = __builtin_fabs(gmem(bvec)[r]);
        bn = nan_max(v, bn);
    }
    const double bnorm = block_max(L, bn);
    double amax = 0.0;
    const double anorm = copy_in(A, lda, W, ldw, n, L, amax);
    for (int r = t; r < n; r += BT) L.perm[r] = r;
    __syncthreads();
    // Mixed precision: an entry beyond the FP32 range (or a NaN) cannot be
    // factored in FP32 (LAPACK dsgesv's dlag2s check).  Such a matrix is not
    // solved here: info = n + 2, iters = -2, x untouched; the plain FP64
    // solver takes it.  The test is on a wave-uniform value.
    if (MIXED && !(amax <= 3.4028234663852886e38)) {
        if (t == 0) {
            G.info[bi] = n + 2;
            G.iters[bi] = -2;
            G.berr[bi] = __builtin_inf();
            G.state[bi] = 1;
        }
        return;
    }

    int zero = 0;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = n - k0 < NB ? n - k0 : NB;
        const int z = panel<T>(W, ldw, n, k0, nb, L);
        if (zero == 0) zero = z;
        swap_trsm<T>(W, ldw, n, k0, nb, L);
        if (k0 + nb < n) {
            if constexpr (MIXED) {
                if (n >= 256) trailing_bf16x6(W, ldw, n, k0, L);
                else trailing_f32_valu(W, ldw, n, k0, L);
            } else {
                trailing_f64(W, ldw, n, k0, L);
            }
        }
    }
    int* perm = G.perm + (long long)bi * ldw;
    for (int r = t; r < n; r += BT) perm[r] = L.perm[r];
    if (t == 0) {
        G.norms[2 * bi] = anorm;
        G.norms[2 * bi + 1] = bnorm;
        G.zero[bi] = zero;
        G.state[bi] = 0;
    }
}

// One step of the solve for matrix bi.  Step 0: x = (LU)^-1 P b.  Step
// it > 0: d = (LU)^-1 P r, x += d (r from the previous step).  Then
// r = b - A x and eta = ||r|| / (||A|| ||x|| + ||b||), all in FP64; the
// matrix is finished when eta <= tol * n, at step max_iter, or at once
// without refinement (the plain FP64 path).
template <typename T, bool MIXED>
__device__ __forceinline__ void gesv_step(const GesvArgs& G, const Lds<T>& L, int bi, int n, int lda, int it)
{
    const int ldw = G.ldw;
    const int t = threadIdx.x;
    const double* A = G.a[bi];
    const double* bvec = G.b[bi];
    double* xo = G.x[bi];
    const T* W = (const T*)(G.ws + (long long)bi * G.wstride);
    const int* perm = G.perm + (long long)bi * ldw;
    double* rs = G.resid + (long long)bi * ldw;
    for (int r = t; r < n; r += BT) {
        const int pr = gmem(perm)[r];
        L.yv[r] = (T)(it == 0 ? gmem(bvec)[pr] : gmem(rs)[pr]);
    }
    __syncthreads();
    lu_solve<T>(W, ldw, n, L);
    double xm = 0.0;
    for (int r = t; r < n; r += BT) {
        const double v = it == 0 ? (double)L.yv[r] : gmem(xo)[r] + (double)L.yv[r];
        L.xv[r] = v;
        gmem_w(xo)[r] = v;
        const double a = __builtin_fabs(v);
        xm = nan_max(a, xm);
    }
    const double xnorm = block_max(L, xm);  // its barriers also publish xv
    const double rnorm = residual<T>(A, lda, bvec, n, L);
    for (int r = t; r < n; r += BT) gmem_w(rs)[r] = L.rv[r];
    const double anorm = G.norms[2 * bi], bnorm = G.norms[2 * bi + 1];
    const double den = anorm * xnorm + bnorm;
    // NaN norms (a NaN in A, b or x) give a NaN eta; a zero denominator
    // (A = 0 and b = 0) gives 0 for r = 0
    const double eta = uniform((den != den || rnorm != rnorm) ? __builtin_nan("")
                                                              : (den > 0.0 ? rnorm / den : (rnorm > 0.0 ? 1.0 : 0.0)));
    // A NaN backward error stops the refinement and is a failure, like one
    // that stays above the tolerance; for the plain FP64 solve a backward
    // error that is not finite is one.
    const bool done = !MIXED || !(eta > G.tol_per_n * n) || it >= G.max_iter;
    if (t == 0) {
        G.berr[bi] = eta;
        if (MIXED) G.iters[bi] = it;
        if (done) {
            const int zero = G.zero[bi];
            const bool bad = MIXED ? !(eta <= G.tol_per_n * n) : !(eta <= __builtin_huge_val());
            G.info[bi] = zero ? zero : (bad ? n + 1 : 0);
            G.state[bi] = 1;
        } else {
            atomicAdd(&G.active[it + 1], 1);
        }
    }
}

// The persistent loops have a single back edge and every branch around a
// barrier is on a wave-uniform value.
#ifndef MFH_GESV_WAVES
#define MFH_GESV_WAVES 2
#endif
template <typename T, bool MIXED>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(MFH_GESV_WAVES))) void k_gesv(GesvArgs G)
{
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const Lds<T> L = carve<T>(lds_raw, G.ldw);
    const int t = threadIdx.x;
    for (;;) {
        if (t == 0) {
            const int k = atomicAdd(G.counter, 1);
            *L.flag = k < G.count ? G.order[k] : G.count;
        }
        __syncthreads();
        const int bi = __builtin_amdgcn_readfirstlane(*L.flag);
        __syncthreads();
        if (bi >= G.count) break;
        const int n = __builtin_amdgcn_readfirstlane(G.n[bi]);
        const int lda = __builtin_amdgcn_readfirstlane(G.lda[bi]);
        int bad = 0;
        if (n < 1 || n > G.max_n) bad = -3;
        else if (G.a[bi] == nullptr) bad = -4;
        else if (lda < n || (long long)lda * n > MFH_MAX_MATRIX_ELEMS) bad = -5;
        else if (!G.factor_only && G.b[bi] == nullptr) bad = -6;
        else if (!G.factor_only && G.x[bi] == nullptr) bad = -7;
        else if (G.factor_only && (G.lu_out[bi] == nullptr || G.piv_out[bi] == nullptr)) bad = -6;
        bad = __builtin_amdgcn_readfirstlane(bad);
        if (bad == 0) gesv_factor<T, MIXED>(G, L, bi, n, lda);
        else if (t == 0) {
            G.info[bi] = bad;
            G.state[bi] = -1;
        }
        __syncthreads();
    }
}

// LDS of the solve launches: xv, rv [ldw] double, yv [ldw] T, the group
// partial sums of the residual [BT] double, the block-reduction slots and
// the staged diagonal block of the triangular solves.
template <typename T>
__host__ __device__ size_t step_lds_bytes(int ldw)
{
    constexpr int BS = Prec<T>::BS;
    return (size_t)ldw * (2 * sizeof(double) + sizeof(T)) + BT * sizeof(double) + 2 * NWAVE * sizeof(double) + 64 +
           (size_t)BS * (BS + 1) * sizeof(T) + 16;
}

template <typename T>
__device__ __forceinline__ Lds<T> carve_step(char* base, int ldw)
{
    Lds<T> L{};
    size_t s = 0;
    L.xv = (double*)(base + s);
    s += (size_t)ldw * sizeof(double);
    L.rv = (double*)(base + s);
    s += (size_t)ldw * sizeof(double);
    L.ut = (T*)(base + s);  // residual group sums
    s += BT * sizeof(double);
    L.rkey = (double*)(base + s);
    s += 2 * NWAVE * sizeof(double);
    L.yv = (T*)(base + s);
    s += (size_t)ldw * sizeof(T);
    s = (s + 15) & ~(size_t)15;
    L.diag = (T*)(base + s);
    return L;
}

// Solve step `it` for every matrix still active; workgroups take the
// matrices largest first through the step's work counter, so the launch
// ends with the small ones (a fixed stride over the batch order left some
// workgroups with several times the work of others).  The solve and the residual are bound by memory
// latency: this kernel is held to MFH_STEP_WAVES wavefronts per SIMD (80
// VGPRs at 6, no spills), so that three workgroups share a CU and their
// loads overlap.
#ifndef MFH_STEP_WAVES
#define MFH_STEP_WAVES 6
#endif
template <typename T, bool MIXED>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(MFH_STEP_WAVES))) void k_gesv_step(GesvArgs G, int it)
{
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const Lds<T> L = carve_step<T>(lds_raw, G.ldw);
    // a refinement step with nothing left to refine ends at once (the
    // previous step counted the matrices it left active; every workgroup
    // reads the same count)
    if (it > 0 && __builtin_amdgcn_readfirstlane(*(volatile int*)&G.active[it]) == 0) return;
    int* slot = (int*)L.rkey;  // free between matrices
    for (;;) {
        if (threadIdx.x == 0) {
            // the matrix at the next position of the order (one that has
            // already finished is skipped below)
            const int k = atomicAdd(&G.stepc[it], 1);
            *slot = k < G.count ? G.order[k] : G.count;
        }
        __syncthreads();
        const int bi = __builtin_amdgcn_readfirstlane(*slot);
        __syncthreads();
        if (bi >= G.count) break;
        const int active = __builtin_amdgcn_readfirstlane(G.state[bi]) == 0;
        if (active) {
            const int n = __builtin_amdgcn_readfirstlane(G.n[bi]);
            const int lda = __builtin_amdgcn_readfirstlane(G.lda[bi]);
            gesv_step<T, MIXED>(G, L, bi, n, lda, it);
        }
        __syncthreads();
    }
}

// Factor-only mode: writes the FP64 factors from the row-major workspace W
// back to the caller's column-major matrix (through 32 x 33 LDS tiles, both
// sides coalesced), the row order to piv and the first zero pivot to info.
__global__ __launch_bounds__(BT) void k_gesv_export(GesvArgs G)
{
    __shared__ double tile[32 * 33];
    const int t = threadIdx.x;
    for (int bi = blockIdx.x; bi < G.count; bi += gridDim.x) {
        const int st = __builtin_amdgcn_readfirstlane(G.state[bi]);
        if (st == 0) {
            const int n = __builtin_amdgcn_readfirstlane(G.n[bi]);
            const int lda = __builtin_amdgcn_readfirstlane(G.lda[bi]);
            const double* W = (const double*)(G.ws + (long long)bi * G.wstride);
            double* out = G.lu_out[bi];
            for (int r0 = 0; r0 < n; r0 += 32)
                for (int c0 = 0; c0 < n; c0 += 32) {
                    // read rows of W (columns fastest), write columns of out
                    for (int e = t; e < 32 * 32; e += BT) {
                        const int rr = e >> 5, cc = e & 31;
                        const int r = r0 + rr, c = c0 + cc;
                        tile[rr * 33 + cc] = (r < n && c < n) ? gmem(W)[(size_t)r * G.ldw + c] : 0.0;
                    }
                    __syncthreads();
                    for (int e = t; e < 32 * 32; e += BT) {
                        const int cc = e >> 5, rr = e & 31;
                        const int r = r0 + rr, c = c0 + cc;
                        if (r < n && c < n) gmem_w(out)[r + (size_t)c * lda] = tile[rr * 33 + cc];
                    }
                    __syncthreads();
                }
            const int* perm = G.perm + (long long)bi * G.ldw;
            int* piv = G.piv_out[bi];
            for (int r = t; r < n; r += BT) gmem_w(piv)[r] = gmem(perm)[r];
            if (t == 0) G.info[bi] = G.zero[bi];
        }
        __syncthreads();
    }
}

template <typename T, bool MIXED>
int launch_gesv(mfh_handle* h, GesvArgs G)
{
    const size_t lds = gesv_lds_bytes<T>(G.ldw);
    MFH_HIP_CHECK(hipFuncSetAttribute((const void*)k_gesv<T, MIXED>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gesv<T, MIXED>, BT, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > G.count) grid = G.count;
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(256), 0, h->stream, G.counter, kGesvHeaderInts);
    // largest first: a counting sort of the ids by n / 8
    static_assert(MfhOrder<MFH_BLOCKED_MAX_N>::keys <= kGesvCursorOff - kGesvHistOff, "order histogram");
    mfh_launch_order<MFH_BLOCKED_MAX_N>(h->stream, G.n, G.count, G.counter + kGesvHistOff,
                                        G.counter + kGesvCursorOff, G.order);
    MFH_HIP_CHECK(hipGetLastError());
    int st = mfh_timing_begin(h, MIXED ? MFH_KERNEL_GESV_MIXED : MFH_KERNEL_GESV_F64, G.max_n);
    if (st) return st;
    hipLaunchKernelGGL((k_gesv<T, MIXED>), dim3((unsigned)grid), dim3(BT), lds, h->stream, G);
    MFH_HIP_CHECK(hipGetLastError());
    st = mfh_timing_end(h);
    if (st) return st;

    if (G.factor_only) {
        per_cu = 0;
        MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gesv_export, BT, 0));
        if (per_cu < 1) per_cu = 1;
        grid = (long)per_cu * h->num_cus;
        if (grid > G.count) grid = G.count;
        hipLaunchKernelGGL(k_gesv_export, dim3((unsigned)grid), dim3(BT), 0, h->stream, G);
        MFH_HIP_CHECK(hipGetLastError());
        return MFH_OK;
    }
    const size_t slds = step_lds_bytes<T>(G.ldw);
    MFH_HIP_CHECK(hipFuncSetAttribute((const void*)k_gesv_step<T, MIXED>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)slds));
    per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gesv_step<T, MIXED>, BT, slds));
    if (per_cu < 1) per_cu = 1;
    grid = (long)per_cu * h->num_cus;
    if (grid > G.count) grid = G.count;
    // step 0 solves; steps 1..max_iter refine what is still active
    const int steps = MIXED ? 1 + G.max_iter : 1;
    for (int it = 0; it < steps; ++it) {
        st = mfh_timing_begin(h, MFH_KERNEL_GESV_STEP, G.max_n);
        if (st) return st;
        hipLaunchKernelGGL((k_gesv_step<T, MIXED>), dim3((unsigned)grid), dim3(BT), slds, h->stream, G, it);
        MFH_HIP_CHECK(hipGetLastError());
        st = mfh_timing_end(h);
        if (st) return st;
    }
    return MFH_OK;
}

}  // namespace

namespace {
struct GesvLayout {
    size_t per, w, perm, resid, norms, zero, state, order, total;
};

GesvLayout gesv_layout(int count, int max_n, bool mixed)
{
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t ldw = (size_t)((max_n + 7) & ~7);
    GesvLayout g;
    g.per = up(ldw * max_n * (mixed ? 4 : 8));
    g.w = sizeof(int) * kGesvHeaderInts;  // counters and the order histogram
    g.perm = g.w + g.per * (size_t)count;
    g.resid = g.perm + up(sizeof(int) * ldw * count);
    g.norms = g.resid + up(sizeof(double) * ldw * count);
    g.zero = g.norms + up(sizeof(double) * 2 * count);
    g.state = g.zero + up(sizeof(int) * count);
    g.order = g.state + up(sizeof(int) * count);
    g.total = g.order + up(sizeof(int) * count);
    return g;
}
}  // namespace

size_t mfh_gesv_ws_bytes(int count, int max_n, bool mixed) { return gesv_layout(count, max_n, mixed).total; }

static int launch_gesv_common(mfh_handle* h, int count, const int* n, const double* const* a, const int* lda,
                              const double* const* b, double* const* x, int max_iter, double tol_per_n, int* iters,
                              double* berr, int* info, int max_n, bool mixed, double* const* lu_out,
                              int* const* piv_out)
{
    const int ldw = (max_n + 7) & ~7;
    const GesvLayout lay = gesv_layout(count, max_n, mixed);
    int st = mfh_ws_reserve(h, lay.total);
    if (st) return st;
    GesvArgs G;
    G.n = n;
    G.a = a;
    G.lda = lda;
    G.b = b;
    G.x = x;
    G.iters = iters;
    G.berr = berr;
    G.info = info;
    G.count = count;
    G.max_n = max_n;
    G.max_iter = max_iter;
    G.tol_per_n = tol_per_n;
    char* base = (char*)h->ws;
    G.counter = (int*)base;
    G.active = (int*)base + 1;
    G.stepc = (int*)base + kGesvStepOff;
    G.order = (int*)(base + lay.order);
    G.ws = base + lay.w;
    G.wstride = (long long)lay.per;
    G.ldw = ldw;
    G.perm = (int*)(base + lay.perm);
    G.resid = (double*)(base + lay.resid);
    G.norms = (double*)(base + lay.norms);
    G.zero = (int*)(base + lay.zero);
    G.state = (int*)(base + lay.state);
    G.factor_only = lu_out != nullptr;
    G.lu_out = lu_out;
    G.piv_out = piv_out;
    return mixed ? launch_gesv<float, true>(h, G) : launch_gesv<double, false>(h, G);
}

int mfh_launch_gesv_mixed(mfh_handle* h, int count, const int* n, const double* const* a, const int* lda,
                          const double* const* b, double* const* x, int max_iter, double tol_per_n, int* iters,
                          double* berr, int* info, int max_n, bool mixed)
{
    return launch_gesv_common(h, count, n, a, lda, b, x, max_iter, tol_per_n, iters, berr, info, max_n, mixed, nullptr,
                              nullptr);
}

int mfh_launch_getrf_blocked(mfh_handle* h, int count, const int* n, double* const* a, const int* lda,
                             int* const* piv, int* info, int max_n)
{
    return launch_gesv_common(h, count, n, (const double* const*)a, lda, nullptr, nullptr, 0, 1.0, nullptr, nullptr,
                              info, max_n, false, a, piv);
}
