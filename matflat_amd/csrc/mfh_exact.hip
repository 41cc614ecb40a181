// mfh_exact.hip -- batched LU, Cholesky, SPD solve, triangular solve and LU
// solve for gfx950, reproducing MatFlat's FP64 operation order exactly.
//
// Design (MI355X):
//  * One workgroup per matrix, the whole matrix staged in LDS (n <= 128, at
//    most 133 KB of the CU's 160 KB); one thread per row, so a workgroup is
//    64 * ceil(rows / 64) threads (1-3 wavefronts).
//  * Persistent kernels: matrices are bucketed by size on the device (16-wide
//    buckets), and each bucket's kernel is launched with as many workgroups
//    as fit on the 256 CUs; workgroups pop matrix ids from the bucket's
//    atomic queue until it is drained.  The LDS allocation and thread count
//    therefore follow the bucket, not the largest matrix of the batch.
//  * The factorizations are scheduled column by column (Crout order): at
//    step k every row computes its dot product for column k in parallel.
//    Each output element is still exactly one Internals.Dot of MatFlat's
//    length and pairing, fed with the same operands, so the results are
//    bit-identical to MatFlat's row/column-sequential loops.
//  * Pivot search is a wavefront (|w|, index) max-reduction with MatFlat's
//    first-maximum and NaN rules (mfh_pivot_key).
#include "mfh_device.h"
#include "mfh_internal.h"

#include <climits>

#pragma clang fp contract(off)

namespace {

constexpr int kBucketBound(int b) { return MFH_BUCKET_STEP * (b + 1); }

__host__ __device__ inline int mfh_bucket_of(int key) { return (key + MFH_BUCKET_STEP - 1) / MFH_BUCKET_STEP - 1; }

size_t exact_lds_bytes(int nb) { return (size_t)nb * nb * 8 + (size_t)nb * 8 + (size_t)nb * 4 + 64; }

// Workspace: counts[NB] heads[NB] items[NB * count]
struct BucketWs {
    int* counts;
    int* heads;
    int* items;
};

BucketWs bucket_ws(mfh_handle* h, int count)
{
    int* base = (int*)h->ws;
    return BucketWs{base, base + MFH_NBUCKETS, base + 2 * MFH_NBUCKETS};
}

size_t bucket_ws_bytes(int count) { return sizeof(int) * (2 * MFH_NBUCKETS + (size_t)MFH_NBUCKETS * count); }

__global__ void k_reset(int* p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

__device__ inline void bucket_push(BucketWs ws, int count, int key, int i)
{
    int b = mfh_bucket_of(key);
    int pos = atomicAdd(&ws.counts[b], 1);
    ws.items[(size_t)b * count + pos] = i;
}

// ------------------------------------------------------------------ LU
__global__ void k_bucketize_getrf(MfhGetrfBatch B, int count, BucketWs ws, int max_m, int max_n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int m = B.m ? B.m[i] : B.m0;
    int n = B.n ? B.n[i] : B.n0;
    int lda = B.lda ? B.lda[i] : B.lda0;
    double* a = B.a ? B.a[i] : B.a0;
    int* piv = B.piv ? B.piv[i] : B.piv0;
    int bad = 0;
    if (m < 1 || m > max_m || m > MFH_EXACT_MAX_N) bad = -3;
    else if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -4;
    else if (a == nullptr) bad = -5;
    else if (lda < m) bad = -6;
    else if (piv == nullptr) bad = -7;
    if (bad) {
        B.info[i] = bad;
        return;
    }
    bucket_push(ws, count, m > n ? m : n, i);
}

template <int T>
__global__ __launch_bounds__(T) void k_getrf_exact(MfhGetrfBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    __shared__ int s_slot;
    __shared__ double s_rkey[T / MFH_WAVE];
    __shared__ int s_ridx[T / MFH_WAVE];
    __shared__ int s_p;
    const int tid = threadIdx.x;
    for (;;) {
        const int b = mfh_pop(q, &s_slot);
        if (b < 0) break;
        const int m = B.m ? B.m[b] : B.m0;
        const int n = B.n ? B.n[b] : B.n0;
        const int lda = B.lda ? B.lda[b] : B.lda0;
        double* A = B.a ? B.a[b] : B.a0 + (size_t)b * B.stride_a;
        int* P = B.piv ? B.piv[b] : B.piv0 + (size_t)b * B.stride_piv;
        double* L = lds;  // m x n, leading dimension m
        int* pv = (int*)(lds + m * n);

        for (int e = tid; e < m * n; e += T) {
            int j = e / m;
            int i = e - j * m;
            L[e] = A[i + (size_t)j * lda];
        }
        for (int i = tid; i < m; i += T) pv[i] = i;
        __syncthreads();

        int sign = 1;
        const int kmax = m < n ? m : n;
        for (int k = 0; k < kmax; ++k) {
            // Column k below (and at) the diagonal: Factorization.Lu.cs:315-320
            // for rows i >= j, a dot of length j.
            const int i = k + tid;
            const bool act = i < m;
            double w = 0.0;
            if (act) {
                w = L[i + k * m] - mfh_dot_exact(k, L + i, m, L + k * m, 1);
                L[i + k * m] = w;
            }
            // Pivot (Lu.cs:323-330).
            double key = act ? mfh_pivot_key(w) : -2.0;
            int idx = act ? i : INT_MAX;
            mfh_wave_argmax(key, idx);
            if ((tid & (MFH_WAVE - 1)) == 0) {
                s_rkey[tid / MFH_WAVE] = key;
                s_ridx[tid / MFH_WAVE] = idx;
            }
            __syncthreads();
            if (tid == 0) {
                double kk = s_rkey[0];
                int ii = s_ridx[0];
                for (int wv = 1; wv < T / MFH_WAVE; ++wv) mfh_pivot_combine(kk, ii, s_rkey[wv], s_ridx[wv]);
                s_p = __builtin_isnan(w) ? k : ii;  // thread 0 holds w_k
            }
            __syncthreads();
            const int p = s_p;
            if (p != k) {
                // Internals.SwapRows over all n columns (Lu.cs:332-337).
                for (int j = tid; j < n; j += T) {
                    double t = L[k + j * m];
                    L[k + j * m] = L[p + j * m];
                    L[p + j * m] = t;
                }
                if (tid == 0) {
                    int t = pv[k];
                    pv[k] = pv[p];
                    pv[p] = t;
                }
                sign = -sign;
                __syncthreads();
            }
            // Multipliers by true division (Lu.cs:339-344).
            const double diag = L[k + k * m];
            if (act && i > k && diag != 0.0) L[i + k * m] = L[i + k * m] / diag;
            // Row k of U for columns j > k: the dot of length k that the
            // reference computes when it reaches column j (Lu.cs:315-320, i < j).
            const int j = k + 1 + tid;
            if (j < n) L[k + j * m] = L[k + j * m] - mfh_dot_exact(k, L + k, m, L + j * m, 1);
            __syncthreads();
        }

        for (int e = tid; e < m * n; e += T) {
            int j = e / m;
            int i = e - j * m;
            A[i + (size_t)j * lda] = L[e];
        }
        for (int i = tid; i < m; i += T) P[i] = pv[i];
        if (tid == 0) {
            B.sign[b] = sign;
            B.info[b] = 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ Cholesky / posv
__global__ void k_bucketize_chol(MfhCholBatch B, int count, BucketWs ws, int max_n, int posv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int n = B.n[i];
    int bad = 0;
    if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -3;
    else if (B.a[i] == nullptr) bad = -4;
    else if (B.lda[i] < n) bad = -5;
    else if (posv && B.x[i] == nullptr) bad = -6;
    if (bad) {
        B.info[i] = bad;
        return;
    }
    bucket_push(ws, count, posv ? n + 1 : n, i);
}

// Factorization.Cholesky.cs:114-168, scheduled by columns: at step k every
// row j > k computes L[j,k] = (A[k,j] - Dot(k, L[j,0:k], L[k,0:k])) / L[k,k]
// and adds t*t to its own running sum (the reference's s, accumulated in
// the same k order).  The thread of row k+1 then finishes its diagonal in
// the same step.  With POSV the right-hand side is carried as an extra row
// n (b in column n of the LDS copy): its "L[n,k]" is exactly the forward
// substitution of SolveTriangular(Lower, NoTrans) (Blas.SolveTriangular.cs:
// 216-222), after which the Lower/Trans sweep (:224-231) runs with one
// thread per unknown.
template <int T, bool POSV>
__global__ __launch_bounds__(T) void k_potrf_exact(MfhCholBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    __shared__ int s_slot;
    __shared__ int s_fail;
    const int tid = threadIdx.x;
    for (;;) {
        const int b = mfh_pop(q, &s_slot);
        if (b < 0) break;
        const int n = B.n[b];
        const int lda = B.lda[b];
        double* A = B.a[b];
        double* X = POSV ? B.x[b] : nullptr;
        const int rows = POSV ? n + 1 : n;
        const int ls = rows;
        double* S = lds;
        double* xs = lds + ls * rows;

        // Upper triangle only (the lower one is never read).
        for (int j = 0; j < n; ++j)
            for (int i = tid; i <= j; i += T) S[i + j * ls] = A[i + (size_t)j * lda];
        if (POSV)
            for (int i = tid; i < n; i += T) S[i + n * ls] = X[i];
        if (tid == 0) {
            s_fail = 0;
            double d = S[0] - 0.0;
            if (d > 0) S[0] = __builtin_sqrt(d);
            else s_fail = 1;
        }
        double sj = 0.0;
        __syncthreads();
        int fail = s_fail;
        const int steps = POSV ? n : n - 1;
        for (int k = 0; k < steps && !fail; ++k) {
            const int j = tid;
            if (j > k && j < rows) {
                double t = (S[k + j * ls] - mfh_dot_exact(k, S + j, ls, S + k, ls)) / S[k + k * ls];
                S[j + k * ls] = t;
                if (j < n) {
                    sj = sj + t * t;
                    if (j == k + 1) {
                        double d = S[j + j * ls] - sj;
                        if (d > 0) S[j + j * ls] = __builtin_sqrt(d);
                        else s_fail = j + 1;
                    }
                }
            }
            __syncthreads();
            fail = s_fail;
        }

        if (fail) {
            // The reference throws at row fail-1 after writing rows
            // 0..fail-2 completely and row fail-1 left of the diagonal.
            const int jf = fail - 1;
            for (int c = 0; c < jf; ++c)
                for (int i = c + tid; i <= jf; i += T) {
                    if (i == c && c == jf) continue;
                    A[i + (size_t)c * lda] = S[i + c * ls];
                }
            if (tid == 0) B.info[b] = fail;
        } else {
            for (int c = 0; c < n; ++c)
                for (int i = tid; i < n; i += T) A[i + (size_t)c * lda] = (i < c) ? 0.0 : S[i + c * ls];
            if (POSV) {
                double xr = (tid < n) ? S[n + tid * ls] : 0.0;
                if (tid == n - 1) {
                    xr = xr / S[(n - 1) + (n - 1) * ls];
                    xs[n - 1] = xr;
                }
                __syncthreads();
                for (int j = n - 1; j >= 1; --j) {
                    if (tid < j) {
                        xr = xr - S[j + tid * ls] * xs[j];
                        if (tid == j - 1) {
                            xr = xr / S[tid + tid * ls];
                            xs[tid] = xr;
                        }
                    }
                    __syncthreads();
                }
                if (tid < n) X[tid] = xr;
            }
            if (tid == 0) B.info[b] = 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ triangular solve
__global__ void k_bucketize_trsv(MfhTrsvBatch B, int count, BucketWs ws, int max_n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int n = B.n[i];
    int bad = 0;
    if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -6;
    else if (B.a[i] == nullptr) bad = -7;
    else if (B.lda[i] < n) bad = -8;
    else if (B.x[i] == nullptr) bad = -9;
    else if (B.incx[i] < 1) bad = -10;
    if (bad) {
        B.info[i] = bad;
        return;
    }
    bucket_push(ws, count, n, i);
}

// Progressive form of a dot-product substitution: thread i owns unknown i
// and folds the product with x_c into its running Internals.Dot as soon as
// x_c is final, keeping the reference's pairing (length i: odd -> x0*y0 then
// pairs (1,2),(3,4)...; even -> pairs (0,1),(2,3)...).
__device__ __forceinline__ void dot_fold(int i, int c, double p, double& s, double& pend)
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
__global__ __launch_bounds__(T) void k_trsv_exact(MfhTrsvBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    __shared__ int s_slot;
    const int tid = threadIdx.x;
    const bool upper = B.uplo == MFH_UPPER;
    const bool trans = B.trans == MFH_TRANS;
    const bool unit = B.unit != 0;
    for (;;) {
        const int b = mfh_pop(q, &s_slot);
        if (b < 0) break;
        const int n = B.n[b];
        const int lda = B.lda[b];
        const int incx = B.incx[b];
        const double* A = B.a[b];
        double* X = B.x[b];
        double* S = lds;
        double* xs = lds + n * n;
        for (int j = 0; j < n; ++j) {
            if (upper)
                for (int i = tid; i <= j; i += T) S[i + j * n] = A[i + (size_t)j * lda];
            else
                for (int i = j + tid; i < n; i += T) S[i + j * n] = A[i + (size_t)j * lda];
        }
        double xr = (tid < n) ? X[(size_t)tid * incx] : 0.0;
        __syncthreads();
        if (upper == trans) {
            // Dot forms: Lower/NoTrans (Blas.SolveTriangular.cs:216-222) reads
            // row i of L, Upper/Trans (:201-207) reads column i of U.
            double s = 0.0, pend = 0.0;
            if (tid == 0) {
                xr = xr - 0.0;
                if (!unit) xr = xr / S[0];
                xs[0] = xr;
            }
            __syncthreads();
            for (int c = 0; c < n - 1; ++c) {
                const int i = tid;
                if (i > c && i < n) {
                    double aic = upper ? S[c + i * n] : S[i + c * n];
                    dot_fold(i, c, aic * xs[c], s, pend);
                    if (i == c + 1) {
                        xr = xr - s;
                        if (!unit) xr = xr / S[i + i * n];
                        xs[i] = xr;
                    }
                }
                __syncthreads();
            }
        } else {
            // Axpy forms, bottom-up: Upper/NoTrans (:191-199) subtracts column
            // i of U, Lower/Trans (:224-231) row j of L (MulSub :397-421).
            if (tid == n - 1) {
                if (!unit) xr = xr / S[(n - 1) * (n + 1)];
                xs[n - 1] = xr;
            }
            __syncthreads();
            for (int j = n - 1; j >= 1; --j) {
                if (tid < j) {
                    double arj = upper ? S[tid + j * n] : S[j + tid * n];
                    xr = xr - arj * xs[j];
                    if (tid == j - 1) {
                        if (!unit) xr = xr / S[tid + tid * n];
                        xs[tid] = xr;
                    }
                }
                __syncthreads();
            }
        }
        if (tid < n) X[(size_t)tid * incx] = xr;
        if (tid == 0) B.info[b] = 0;
        __syncthreads();
    }
}

// ------------------------------------------------------------ LU solve
__global__ void k_bucketize_getrs(MfhGetrsBatch B, int count, BucketWs ws, int max_n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int n = B.n[i];
    int bad = 0;
    if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -3;
    else if (B.lu[i] == nullptr) bad = -4;
    else if (B.lda[i] < n) bad = -5;
    else if (B.piv[i] == nullptr) bad = -6;
    else if (B.b[i] == nullptr) bad = -7;
    else if (B.x[i] == nullptr) bad = -8;
    if (bad) {
        B.info[i] = bad;
        return;
    }
    bucket_push(ws, count, n, i);
}

// Factorization.LuInverse.cs:169-192 for one column: permute, unit-lower
// dot-form sweep (the reference divides by the 1.0 it stores on the copy's
// diagonal, which is exact), then the Upper/NoTrans axpy sweep.
template <int T>
__global__ __launch_bounds__(T) void k_getrs_exact(MfhGetrsBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    __shared__ int s_slot;
    const int tid = threadIdx.x;
    for (;;) {
        const int b = mfh_pop(q, &s_slot);
        if (b < 0) break;
        const int n = B.n[b];
        const int lda = B.lda[b];
        const double* LU = B.lu[b];
        const int* P = B.piv[b];
        double* S = lds;
        double* xs = lds + n * n;
        for (int e = tid; e < n * n; e += T) {
            int j = e / n;
            int i = e - j * n;
            S[e] = LU[i + (size_t)j * lda];
        }
        double xr = 0.0;
        if (tid < n) {
            int pr = P[tid];
            xr = (pr >= 0 && pr < n) ? B.b[b][pr] : __builtin_nan("");
        }
        __syncthreads();
        double s = 0.0, pend = 0.0;
        if (tid == 0) {
            xr = xr - 0.0;
            xs[0] = xr;
        }
        __syncthreads();
        for (int c = 0; c < n - 1; ++c) {
            const int i = tid;
            if (i > c && i < n) {
                dot_fold(i, c, S[i + c * n] * xs[c], s, pend);
                if (i == c + 1) {
                    xr = xr - s;
                    xs[i] = xr;
                }
            }
            __syncthreads();
        }
        if (tid == n - 1) {
            xr = xr / S[(n - 1) * (n + 1)];
            xs[n - 1] = xr;
        }
        __syncthreads();
        for (int j = n - 1; j >= 1; --j) {
            if (tid < j) {
                xr = xr - S[tid + j * n] * xs[j];
                if (tid == j - 1) {
                    xr = xr / S[tid + tid * n];
                    xs[tid] = xr;
                }
            }
            __syncthreads();
        }
        if (tid < n) B.x[b][tid] = xr;
        if (tid == 0) B.info[b] = 0;
        __syncthreads();
    }
}

template <typename Batch, typename K>
int run_bucket(mfh_handle* h, K kernel, int threads, int nb, int count, const Batch& B, BucketWs ws, int bucket)
{
    const size_t lds = exact_lds_bytes(nb > MFH_EXACT_MAX_N + 1 ? MFH_EXACT_MAX_N + 1 : nb);
    MFH_HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > count) grid = count;
    if (grid < 1) grid = 1;
    MfhQueue q{ws.heads + bucket, ws.counts + bucket, ws.items + (size_t)bucket * count};
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(threads), lds, h->stream, B, q);
    MFH_HIP_CHECK(hipGetLastError());
    return MFH_OK;
}

int threads_for(int nb) { return MFH_WAVE * ((nb + MFH_WAVE - 1) / MFH_WAVE); }

int prepare_buckets(mfh_handle* h, int count, BucketWs* ws)
{
    int st = mfh_ws_reserve(h, bucket_ws_bytes(count));
    if (st != MFH_OK) return st;
    *ws = bucket_ws(h, count);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, ws->counts, 2 * MFH_NBUCKETS);
    MFH_HIP_CHECK(hipGetLastError());
    return MFH_OK;
}

int nbuckets_for(int max_key)
{
    int nb = mfh_bucket_of(max_key) + 1;
    return nb > MFH_NBUCKETS ? MFH_NBUCKETS : nb;
}

}  // namespace

int mfh_launch_getrf(mfh_handle* h, int count, const MfhGetrfBatch& B, int max_m, int max_n)
{
    BucketWs ws;
    int st = prepare_buckets(h, count, &ws);
    if (st) return st;
    hipLaunchKernelGGL(k_bucketize_getrf, dim3((count + 255) / 256), dim3(256), 0, h->stream, B, count, ws,
                       max_m, max_n);
    MFH_HIP_CHECK(hipGetLastError());
    int key = max_m > max_n ? max_m : max_n;
    if (key > MFH_EXACT_MAX_N) key = MFH_EXACT_MAX_N;
    for (int b = 0; b < nbuckets_for(key); ++b) {
        int nb = kBucketBound(b);
        int t = threads_for(nb);
        if (t == 64) st = run_bucket(h, k_getrf_exact<64>, 64, nb, count, B, ws, b);
        else st = run_bucket(h, k_getrf_exact<128>, 128, nb, count, B, ws, b);
        if (st) return st;
    }
    return MFH_OK;
}

int mfh_launch_potrf(mfh_handle* h, int count, const MfhCholBatch& B, int max_n, bool posv)
{
    BucketWs ws;
    int st = prepare_buckets(h, count, &ws);
    if (st) return st;
    hipLaunchKernelGGL(k_bucketize_chol, dim3((count + 255) / 256), dim3(256), 0, h->stream, B, count, ws,
                       max_n, posv ? 1 : 0);
    MFH_HIP_CHECK(hipGetLastError());
    int n_cap = max_n > MFH_EXACT_MAX_N ? MFH_EXACT_MAX_N : max_n;
    int key = posv ? n_cap + 1 : n_cap;
    for (int b = 0; b < nbuckets_for(key); ++b) {
        int nb = kBucketBound(b);
        int t = threads_for(nb);
        if (posv) {
            if (t == 64) st = run_bucket(h, k_potrf_exact<64, true>, 64, nb, count, B, ws, b);
            else if (t == 128) st = run_bucket(h, k_potrf_exact<128, true>, 128, nb, count, B, ws, b);
            else st = run_bucket(h, k_potrf_exact<192, true>, 192, nb, count, B, ws, b);
        } else {
            if (t == 64) st = run_bucket(h, k_potrf_exact<64, false>, 64, nb, count, B, ws, b);
            else st = run_bucket(h, k_potrf_exact<128, false>, 128, nb, count, B, ws, b);
        }
        if (st) return st;
    }
    return MFH_OK;
}

int mfh_launch_trsv(mfh_handle* h, int count, const MfhTrsvBatch& B, int max_n)
{
    BucketWs ws;
    int st = prepare_buckets(h, count, &ws);
    if (st) return st;
    hipLaunchKernelGGL(k_bucketize_trsv, dim3((count + 255) / 256), dim3(256), 0, h->stream, B, count, ws, max_n);
    MFH_HIP_CHECK(hipGetLastError());
    int key = max_n > MFH_EXACT_MAX_N ? MFH_EXACT_MAX_N : max_n;
    for (int b = 0; b < nbuckets_for(key); ++b) {
        int nb = kBucketBound(b);
        int t = threads_for(nb);
        if (t == 64) st = run_bucket(h, k_trsv_exact<64>, 64, nb, count, B, ws, b);
        else st = run_bucket(h, k_trsv_exact<128>, 128, nb, count, B, ws, b);
        if (st) return st;
    }
    return MFH_OK;
}

int mfh_launch_getrs(mfh_handle* h, int count, const MfhGetrsBatch& B, int max_n)
{
    BucketWs ws;
    int st = prepare_buckets(h, count, &ws);
    if (st) return st;
    hipLaunchKernelGGL(k_bucketize_getrs, dim3((count + 255) / 256), dim3(256), 0, h->stream, B, count, ws, max_n);
    MFH_HIP_CHECK(hipGetLastError());
    int key = max_n > MFH_EXACT_MAX_N ? MFH_EXACT_MAX_N : max_n;
    for (int b = 0; b < nbuckets_for(key); ++b) {
        int nb = kBucketBound(b);
        int t = threads_for(nb);
        if (t == 64) st = run_bucket(h, k_getrs_exact<64>, 64, nb, count, B, ws, b);
        else st = run_bucket(h, k_getrs_exact<128>, 128, nb, count, B, ws, b);
        if (st) return st;
    }
    return MFH_OK;
}
