// mfh_device.h -- device helpers shared by the gfx950 kernels.
//
// The "exact" helpers reproduce MatFlat's floating-point operation order
// one IEEE binary64 operation at a time; every file that includes them is
// compiled with -ffp-contract=off and the helpers carry the pragma as well,
// so no multiply-add is ever fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MFH_WAVE 64

// Internals.Dot (MatFlat/Internals.cs:265-292): odd length starts with
// x0*y0, even length with 0.0; then sum = sum + (x_k*y_k + x_{k+1}*y_{k+1}).
// W (even) products per step: their loads and products do not depend on s,
// so they are issued together and only the W/2 additions into s stay
// sequential, in MatFlat's order.
template <int W>
__device__ __forceinline__ double mfh_dot_exact_w(int len, const double* x, int incx, const double* y, int incy)
{
#pragma clang fp contract(off)
    static_assert(W % 2 == 0, "products are added in pairs");
    double s;
    int c;
    if (len & 1) {
        s = x[0] * y[0];
        c = 1;
    } else {
        s = 0.0;
        c = 0;
    }
    for (; c + W <= len; c += W) {
        double p[W];
#pragma unroll
        for (int u = 0; u < W; ++u) p[u] = x[(c + u) * incx] * y[(c + u) * incy];
#pragma unroll
        for (int u = 0; u < W; u += 2) s = s + (p[u] + p[u + 1]);
    }
    if constexpr (W > 8) {
        for (; c + 8 <= len; c += 8) {
            double p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] = x[(c + u) * incx] * y[(c + u) * incy];
#pragma unroll
            for (int u = 0; u < 8; u += 2) s = s + (p[u] + p[u + 1]);
        }
    }
    for (; c < len; c += 2) {
        double p0 = x[c * incx] * y[c * incy];
        double p1 = x[(c + 1) * incx] * y[(c + 1) * incy];
        s = s + (p0 + p1);
    }
    return s;
}

__device__ __forceinline__ double mfh_dot_exact(int len, const double* x, int incx, const double* y, int incy)
{
    return mfh_dot_exact_w<8>(len, x, incx, y, incy);
}

// Copies the m x n column-major block at src (leading dimension lds_, global
// memory) into LDS at dst (leading dimension ldd), eight loads in flight per
// thread: a loop of one load and one LDS store per iteration waits for every
// load in turn, which made a 64 x 64 copy cost 64 HBM round trips.  TRI
// selects what is read: 0 the whole block, 1 the upper triangle (i <= j),
// 2 the lower triangle (i >= j); entries outside it are neither read nor
// written.  PK stores the triangle packed by columns instead (upper: column
// j holds rows 0..j at j(j+1)/2; lower: column j holds rows j..n-1 at
// j*ldd - j(j-1)/2, with ldd = n).  Indices past the end are clamped to a
// valid element and unused.
template <int T, int TRI = 0, bool PK = false>
__device__ __forceinline__ void stage_block(double* dst, int ldd, const double* src, int lds_, int m, int n, int tid)
{
    typedef const __attribute__((address_space(1))) double* GSrc;
    const int tot = m * n;
    for (int e0 = 0; e0 < tot; e0 += 8 * T) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * T + tid;
            const int ec = e < tot ? e : tot - 1;
            const int j = ec / m;
            const int i = ec - j * m;
            const bool in = TRI == 0 || (TRI == 1 ? i <= j : i >= j);
            v[u] = in ? ((GSrc)src)[i + (size_t)j * lds_] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * T + tid;
            const int j = e / m;
            const int i = e - j * m;
            const bool in = TRI == 0 || (TRI == 1 ? i <= j : i >= j);
            const int d = !PK ? i + j * ldd
                              : (TRI == 1 ? ((j * (j + 1)) >> 1) + i : j * ldd - ((j * (j - 1)) >> 1) + (i - j));
            if (e < tot && in) dst[d] = v[u];
        }
    }
}

// Pivot key of Factorization.Lu.cs:323-330 for a parallel argmax.  The
// reference scans p = j; if |w_i| > |w_p| then p = i.  That picks the first
// maximum of |w|, never a NaN below the diagonal (comparisons with NaN are
// false), and keeps p = j when w_j itself is NaN.  Mapping NaN to -1 (below
// every |w| >= 0) turns the scan into an associative (key, lowest index)
// max-reduction; the w_j-is-NaN case is handled by the caller.
__device__ __forceinline__ double mfh_pivot_key(double w)
{
    return __builtin_isnan(w) ? -1.0 : __builtin_fabs(w);
}

__device__ __forceinline__ void mfh_pivot_combine(double& key, int& idx, double okey, int oidx)
{
    if (okey > key || (okey == key && oidx < idx)) {
        key = okey;
        idx = oidx;
    }
}

// Pivot key of a candidate as an unsigned 64-bit integer: the bits of |w|
// plus one (the bits of a non-negative double order like the values), 0 for
// NaN and for lanes without a candidate.  A NaN then ranks below every
// number, as in MatFlat's strict '>' scan; the w_k-is-NaN case is handled by
// the caller.
__device__ __forceinline__ unsigned long long mfh_pivot_key_u64(bool act, double w)
{
    const unsigned long long bits = (unsigned long long)__double_as_longlong(__builtin_fabs(w));
    return (act && !__builtin_isnan(w)) ? bits + 1 : 0ull;
}

// One DPP step of a wave maximum of unsigned 64-bit values: lanes the
// control or the row mask leave out read 0, the identity of the maximum.
#define MFH_DPP_MAX_U64(v, CTRL, RMASK)                                                              \
    do {                                                                                           \
        const unsigned lo_ = __builtin_amdgcn_update_dpp(0u, (unsigned)(v), CTRL, RMASK, 0xF, false);   \
        const unsigned hi_ = __builtin_amdgcn_update_dpp(0u, (unsigned)((v) >> 32), CTRL, RMASK, 0xF, false); \
        const unsigned long long o_ = ((unsigned long long)hi_ << 32) | lo_;                          \
        (v) = o_ > (v) ? o_ : (v);                                                                 \
    } while (0)

// Wave-wide maximum through DPP row shifts and row broadcasts (no LDS
// traffic, unlike a shuffle); every lane receives it.
__device__ __forceinline__ unsigned long long mfh_wave_max_u64(unsigned long long v)
{
    MFH_DPP_MAX_U64(v, 0x111, 0xF);  // row_shr:1
    MFH_DPP_MAX_U64(v, 0x112, 0xF);  // row_shr:2
    MFH_DPP_MAX_U64(v, 0x114, 0xF);  // row_shr:4
    MFH_DPP_MAX_U64(v, 0x118, 0xF);  // row_shr:8  -> lane 15 of each row: the row's maximum
    MFH_DPP_MAX_U64(v, 0x142, 0xA);  // row_bcast:15 into rows 1 and 3
    MFH_DPP_MAX_U64(v, 0x143, 0xC);  // row_bcast:31 into rows 2 and 3 -> lane 63: the maximum
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, 63);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// The double lane `lane` (wave-uniform) holds.
__device__ __forceinline__ double mfh_readlane_d(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Wave-wide (key, index) max with lowest-index tie break.
__device__ __forceinline__ void mfh_wave_argmax(double& key, int& idx)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double ok = __shfl_xor(key, off, MFH_WAVE);
        int oi = __shfl_xor(idx, off, MFH_WAVE);
        mfh_pivot_combine(key, idx, ok, oi);
    }
}

// Persistent-kernel work lists.  A bucket's items are the matrix ids of one
// contiguous range of the device-side counting sort (largest size first).
// Workgroup w takes items w, w + grid, w + 2*grid, ...: no shared counter,
// so no single-address atomic traffic (which serialises in L2), and
// neighbouring workgroups get neighbouring (similar-size) matrices.
// A queue with items == nullptr is the whole batch in order: item idx is
// matrix idx for idx < direct (a uniform batch needs no sort).
struct MfhQueue {
    const int* start;  // device: first item of the bucket in items[]
    const int* count;  // device: number of items in the bucket
    const int* items;  // matrix ids sorted by size key, descending
    int direct;        // items == nullptr: the batch size
};

__device__ __forceinline__ int mfh_item(const MfhQueue& q, int idx)
{
    if (q.items == nullptr) return idx < q.direct ? idx : -1;
    return (idx < *q.count) ? q.items[*q.start + idx] : -1;
}
