// mfh_adjoint.hip -- adjoint of the batched SPD solve x = A^-1 b, A = L L^T.
//
// MatFlat has no derivatives; this is the differentiation addition beside
// the plain path.  For loss(x) with xbar = dloss/dx:
//     bbar = A^-1 xbar              (same two sweeps as the forward solve,
//                                    SolveTriangular Lower/NoTrans then
//                                    Lower/Trans, Blas.SolveTriangular.cs:
//                                    216-231, in MatFlat's order)
//     abar = -(bbar x^T + x bbar^T) / 2   (gradient with respect to the
//                                    symmetric matrix the factor came from)
// One workgroup per matrix: the lower triangle of L is staged in LDS (read
// once from HBM), x, xbar and bbar live one per thread, and abar is written
// column by column with coalesced stores -- the kernel moves the bytes the
// result needs and nothing else.
#include "mfh_device.h"
#include "mfh_internal.h"

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ void fold(int i, int c, double p, double& s, double& pend)
{
    if (i & 1) {
        if (c == 0) s = p;
        else if (c & 1) pend = p;
        else s = s + (pend + p);
    } else {
        if ((c & 1) == 0) pend = p;
        else s = s + (pend + p);
    }
}

template <int T>
__global__ __launch_bounds__(T) void k_posv_backward(MfhAdjointBatch B, int count, int max_n)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    // Persistent workgroups walk the batch with a fixed stride; every value
    // that steers a loop or branch around a barrier is wave-uniform.
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const int n = __builtin_amdgcn_readfirstlane(B.n[b]);
        const int lda = __builtin_amdgcn_readfirstlane(B.lda[b]);
        const int ldab = __builtin_amdgcn_readfirstlane(B.ldabar[b]);
        const int bad = __builtin_amdgcn_readfirstlane(
            (n < 1 || n > T || n > max_n || lda < n || ldab < n || B.l[b] == nullptr || B.x[b] == nullptr ||
             B.xbar[b] == nullptr || B.abar[b] == nullptr || B.bbar[b] == nullptr) ? 1 : 0);
        if (bad) {
            if (tid == 0) B.info[b] = -3;
            continue;
        }
        const double* L = B.l[b];
        double* S = lds;
        double* xs = lds + n * n;
        double* xv = xs + n;  // x, shared for the outer product
        stage_block<T, 2>(S, n, L, lda, n, n, tid);
        double br = (tid < n) ? B.xbar[b][tid] : 0.0;
        double xr = (tid < n) ? B.x[b][tid] : 0.0;
        if (tid < n) xv[tid] = xr;
        __syncthreads();
        // Lower/NoTrans dot form.
        double s = 0.0, pend = 0.0;
        if (tid == 0) {
            br = (br - 0.0) / S[0];
            xs[0] = br;
        }
        __syncthreads();
        for (int c = 0; c < n - 1; ++c) {
            const int i = tid;
            if (i > c && i < n) {
                fold(i, c, S[i + c * n] * xs[c], s, pend);
                if (i == c + 1) {
                    br = (br - s) / S[i + i * n];
                    xs[i] = br;
                }
            }
            __syncthreads();
        }
        // Lower/Trans axpy form, bottom-up.
        if (tid == n - 1) {
            br = br / S[(n - 1) * (n + 1)];
            xs[n - 1] = br;
        }
        __syncthreads();
        for (int j = n - 1; j >= 1; --j) {
            if (tid < j) {
                br = br - S[j + tid * n] * xs[j];
                if (tid == j - 1) {
                    br = br / S[tid + tid * n];
                    xs[tid] = br;
                }
            }
            __syncthreads();
        }
        if (tid < n) B.bbar[b][tid] = br;
        double* Ab = B.abar[b];
        for (int j = 0; j < n; ++j) {
            const double bj = xs[j], xj = xv[j];
            for (int i = tid; i < n; i += T) {
                double u = xs[i] * xj;
                double v = xv[i] * bj;
                Ab[i + (size_t)j * ldab] = -0.5 * (u + v);
            }
        }
        if (tid == 0) B.info[b] = 0;
        __syncthreads();
    }
}

template <int T>
int launch_t(mfh_handle* h, int count, const MfhAdjointBatch& B, int max_n)
{
    const size_t lds = (size_t)max_n * max_n * 8 + (size_t)max_n * 16 + 64;
    MFH_HIP_CHECK(hipFuncSetAttribute((const void*)k_posv_backward<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_posv_backward<T>, T, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > count) grid = count;
    hipLaunchKernelGGL(k_posv_backward<T>, dim3((unsigned)grid), dim3(T), lds, h->stream, B, count, max_n);
    MFH_HIP_CHECK(hipGetLastError());
    return MFH_OK;
}

}  // namespace

int mfh_launch_posv_backward(mfh_handle* h, int count, const MfhAdjointBatch& B, int max_n)
{
    if (max_n <= 64) return launch_t<64>(h, count, B, max_n);
    return launch_t<128>(h, count, B, max_n);
}
