// mfh_adjoint.hip -- adjoints of the batched SPD solve x = A^-1 b and of the
// Cholesky factorization A = L L^T.
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

// Adjoint of the Cholesky factorization: Abar = L^-T M L^-1 with M the
// symmetrized Phi(L^T Lbar) / 2 (Phi: lower triangle, diagonal halved; I.
// Murray, arXiv:1602.07527, eq. 10, taken through the symmetric part).
// One workgroup per matrix.  LDS holds L in its lower triangle and Lbar's
// strict lower triangle transposed into the strict upper one (Lbar's
// diagonal beside it), so both inputs cost one matrix of LDS; the work
// matrix W is a second LDS matrix when it fits (WLDS), else the output abar
// itself.  Every value follows the oracle's order exactly:
//   1. one thread per (i >= j): s = sum_{k >= i} L[k][i] Lbar[k][j] from 0.0
//      upward, W[i][j] = W[j][i] = 0.5 s;
//   2. one lane per column x of W: L^T x = m in dot form, j descending,
//      x[j] = (x[j] - Dot(L[j+1.., j], x[j+1..])) / L[j][j] with
//      Internals.Dot's pairing -> L^-T M.  The dot only loads (eight pairs in
//      flight) and each step stores once; the axpy form would store inside
//      its inner loop and wait a full LDS round trip per element;
//   3. one lane per row: the same solve along the row -> L^-T M L^-1;
//   4. abar[i][j] = abar[j][i] = 0.5 (W[i][j] + W[j][i]) for i > j.
template <int T, bool WLDS>
__global__ __launch_bounds__(T) void k_potrf_backward(MfhCholAdjointBatch B, int count, int max_n, int ld)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    double* A1 = lds;
    double* dLb = lds + (size_t)max_n * ld;
    double* Wl = dLb + max_n + 1;  // odd offset keeps the two matrices on different banks
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const int n = __builtin_amdgcn_readfirstlane(B.n[b]);
        const int ldl = __builtin_amdgcn_readfirstlane(B.ldl[b]);
        const int ldlb = __builtin_amdgcn_readfirstlane(B.ldlbar[b]);
        const int ldab = __builtin_amdgcn_readfirstlane(B.ldabar[b]);
        int bad = 0;
        if (n < 1 || n > max_n || n > T) bad = -3;
        else if (B.l[b] == nullptr) bad = -4;
        else if (ldl < n) bad = -5;
        else if (B.lbar[b] == nullptr) bad = -6;
        else if (ldlb < n) bad = -7;
        else if (B.abar[b] == nullptr) bad = -8;
        else if (ldab < n) bad = -9;
        bad = __builtin_amdgcn_readfirstlane(bad);
        if (bad) {
            if (tid == 0) B.info[b] = bad;
            continue;
        }
        const double* L = B.l[b];
        const double* Lb = B.lbar[b];
        double* Ab = B.abar[b];
        double* W = WLDS ? Wl : Ab;
        const int ldw = WLDS ? ld : ldab;
        __syncthreads();  // the previous matrix's reads of LDS are done
        {
            typedef const __attribute__((address_space(1))) double* GSrc;
            const int tot = n * n;
            for (int e0 = 0; e0 < tot; e0 += 8 * T) {  // eight loads of each input in flight per thread
                double lv[8], bv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int e = e0 + u * T + tid;
                    const int ec = e < tot ? e : tot - 1;
                    const int j = ec / n;
                    const int i = ec - j * n;
                    lv[u] = i >= j ? ((GSrc)L)[i + (size_t)j * ldl] : 0.0;
                    bv[u] = i >= j ? ((GSrc)Lb)[i + (size_t)j * ldlb] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int e = e0 + u * T + tid;
                    const int j = e / n;
                    const int i = e - j * n;
                    if (e < tot && i >= j) {
                        A1[i + j * ld] = lv[u];
                        if (i > j) A1[j + i * ld] = bv[u];
                        else dLb[i] = bv[u];
                    }
                }
            }
        }
        __syncthreads();
        for (int e = tid; e < n * n; e += T) {
            const int j = e / n;
            const int i = e - j * n;
            if (i >= j) {
                // s = sum_{k >= i} L[k][i] Lbar[k][j], added in k order; the
                // loads of eight terms go out before their additions
                double s = 0.0;
                int k = i;
                if (k == j) {  // the diagonal term (i == j): Lbar[j][j] sits beside the matrix
                    s = s + A1[k + i * ld] * dLb[j];
                    ++k;
                }
                for (; k + 8 <= n; k += 8) {
                    double p[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) p[u] = A1[k + u + i * ld] * A1[j + (k + u) * ld];
#pragma unroll
                    for (int u = 0; u < 8; ++u) s = s + p[u];
                }
                for (; k < n; ++k) s = s + A1[k + i * ld] * A1[j + k * ld];
                const double m = 0.5 * s;
                W[i + (size_t)j * ldw] = m;
                W[j + (size_t)i * ldw] = m;
            }
        }
        __syncthreads();
        for (int c = tid; c < n; c += T) {
            double* x = W + (size_t)c * ldw;
            for (int jj = n - 1; jj >= 0; --jj) {
                const double d = mfh_dot_exact(n - 1 - jj, A1 + (jj + 1) + jj * ld, 1, x + jj + 1, 1);
                x[jj] = (x[jj] - d) / A1[jj + jj * ld];
            }
        }
        __syncthreads();
        for (int r0 = tid; r0 < n; r0 += T) {
            double* x = W + r0;
            for (int jj = n - 1; jj >= 0; --jj) {
                const double d = mfh_dot_exact(n - 1 - jj, A1 + (jj + 1) + jj * ld, 1, x + (size_t)(jj + 1) * ldw, ldw);
                x[(size_t)jj * ldw] = (x[(size_t)jj * ldw] - d) / A1[jj + jj * ld];
            }
        }
        __syncthreads();
        for (int e = tid; e < n * n; e += T) {
            const int j = e / n;
            const int i = e - j * n;
            if (i > j) {
                const double v = 0.5 * (W[i + (size_t)j * ldw] + W[j + (size_t)i * ldw]);
                Ab[i + (size_t)j * ldab] = v;
                Ab[j + (size_t)i * ldab] = v;
            } else if (i == j && WLDS) {
                Ab[i + (size_t)i * ldab] = W[i + (size_t)i * ldw];
            }
        }
        if (tid == 0) B.info[b] = 0;
    }
}

size_t potrf_backward_lds(int max_n, int ld, bool wlds)
{
    return ((size_t)(wlds ? 2 : 1) * max_n * ld + max_n + 1) * sizeof(double);
}

template <int T, bool WLDS>
int launch_potrf_backward_t(mfh_handle* h, int count, const MfhCholAdjointBatch& B, int max_n, int ld)
{
    const size_t lds = potrf_backward_lds(max_n, ld, WLDS);
    MFH_HIP_CHECK(hipFuncSetAttribute((const void*)k_potrf_backward<T, WLDS>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_potrf_backward<T, WLDS>, T, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > count) grid = count;
    hipLaunchKernelGGL((k_potrf_backward<T, WLDS>), dim3((unsigned)grid), dim3(T), lds, h->stream, B, count, max_n,
                       ld);
    MFH_HIP_CHECK(hipGetLastError());
    return MFH_OK;
}

}  // namespace

int mfh_launch_potrf_backward(mfh_handle* h, int count, const MfhCholAdjointBatch& B, int max_n)
{
    const int ld = max_n | 1;  // odd: lanes walking rows hit distinct LDS banks
    const bool wlds = potrf_backward_lds(max_n, ld, true) <= 160 * 1024;
    if (max_n <= 64) return wlds ? launch_potrf_backward_t<64, true>(h, count, B, max_n, ld)
                                 : launch_potrf_backward_t<64, false>(h, count, B, max_n, ld);
    return wlds ? launch_potrf_backward_t<128, true>(h, count, B, max_n, ld)
                : launch_potrf_backward_t<128, false>(h, count, B, max_n, ld);
}

int mfh_launch_posv_backward(mfh_handle* h, int count, const MfhAdjointBatch& B, int max_n)
{
    if (max_n <= 64) return launch_t<64>(h, count, B, max_n);
    return launch_t<128>(h, count, B, max_n);
}
