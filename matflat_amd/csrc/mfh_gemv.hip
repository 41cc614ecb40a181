// mfh_gemv.hip -- batched matrix-vector product in MatFlat's order.
//
// Blas.MulMatVec (MatFlat/Blas.MulMatVec.cs:138-206) computes every output
// as one Internals.Dot (Internals.cs:265-292): y_i = Dot(n, row i of A, x)
// for NoTrans and y_i = Dot(m, column i of A, x) for Trans.  Each output is
// independent, so lanes take outputs and each lane runs MatFlat's pairwise
// dot sequentially: the results are bit-identical to the CPU path.
//
// Layout and schedule: one wavefront per product at a time, products
// walked with a fixed grid stride (no shared counter).  For NoTrans the 64
// lanes read 64 consecutive rows of one column per load, so every load of A
// is one coalesced 512-byte access; for Trans each lane walks its own
// contiguous column, and the eight loads a lane issues together fall in one
// or two cache lines that the following loads reuse.  x is read by every
// lane at the same address (one cache line serves the wavefront).  The
// bound is HBM: 8 (m n + len(x) + len(y)) bytes per product.
#include "mfh_device.h"
#include "mfh_internal.h"

#pragma clang fp contract(off)

namespace {

// Per-product checks in Blas.MulMatVec.cs:140-183 order; the transpose,
// which the reference rejects only after every other check (:203-205), is
// tested last.  Codes are -(position) in mfh_dgemv_vbatched's argument list.
__device__ __forceinline__ int gemv_check(const MfhGemvBatch& B, int p, int max_m, int max_n)
{
    const int tr = B.trans[p], m = B.m[p], n = B.n[p];
    const int outs = tr == MFH_TRANS ? n : m;
    const int len = tr == MFH_TRANS ? m : n;
    if (m < 1 || m > max_m) return -4;
    if (n < 1 || n > max_n) return -5;
    if (B.a[p] == nullptr) return -6;
    if (B.lda[p] < m || (long long)B.lda[p] * n > MFH_MAX_MATRIX_ELEMS) return -7;
    if (B.x[p] == nullptr) return -8;
    if (B.incx[p] < 1 || (long long)B.incx[p] * len > MFH_MAX_MATRIX_ELEMS) return -9;
    if (B.y[p] == nullptr) return -10;
    if (B.incy[p] < 1 || (long long)B.incy[p] * outs > MFH_MAX_MATRIX_ELEMS) return -11;
    if ((const double*)B.y[p] == B.a[p] || (const double*)B.y[p] == B.x[p]) return -1;
    if (tr != MFH_NOTRANS && tr != MFH_TRANS) return -3;
    return 0;
}

__global__ __launch_bounds__(64) void k_dgemv_exact(MfhGemvBatch B, int count, int max_m, int max_n)
{
    const int lane = threadIdx.x;
    for (int p = blockIdx.x; p < count; p += gridDim.x) {
        const int bad = __builtin_amdgcn_readfirstlane(gemv_check(B, p, max_m, max_n));
        if (lane == 0) B.info[p] = bad;
        if (bad) continue;
        const bool t = __builtin_amdgcn_readfirstlane(B.trans[p]) == MFH_TRANS;
        const int m = __builtin_amdgcn_readfirstlane(B.m[p]);
        const int n = __builtin_amdgcn_readfirstlane(B.n[p]);
        const int lda = __builtin_amdgcn_readfirstlane(B.lda[p]);
        const int incx = __builtin_amdgcn_readfirstlane(B.incx[p]);
        const int incy = __builtin_amdgcn_readfirstlane(B.incy[p]);
        const double* a = B.a[p];
        const double* x = B.x[p];
        double* y = B.y[p];
        if (t) {
            // Blas.MulMatVec.cs:194-199: y_i = Dot(m, a + i*lda, 1, x, incx)
            for (int i = lane; i < n; i += MFH_WAVE) y[i * incy] = mfh_dot_exact_w<16>(m, a + i * lda, 1, x, incx);
        } else {
            // Blas.MulMatVec.cs:186-191: y_i = Dot(n, a + i, lda, x, incx)
            for (int i = lane; i < m; i += MFH_WAVE) y[i * incy] = mfh_dot_exact_w<16>(n, a + i, lda, x, incx);
        }
    }
}

}  // namespace

int mfh_launch_gemv(mfh_handle* h, int count, const MfhGemvBatch& B, int max_m, int max_n)
{
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dgemv_exact, MFH_WAVE, 0));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > count) grid = count;
    int st = mfh_timing_begin(h, MFH_KERNEL_GEMV_EXACT, max_m > max_n ? max_m : max_n);
    if (st) return st;
    hipLaunchKernelGGL(k_dgemv_exact, dim3((unsigned)grid), dim3(MFH_WAVE), 0, h->stream, B, count, max_m, max_n);
    MFH_HIP_CHECK(hipGetLastError());
    return mfh_timing_end(h);
}
