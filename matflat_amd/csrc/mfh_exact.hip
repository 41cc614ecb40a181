// mfh_exact.hip -- batched LU, Cholesky, SPD solve, triangular solve and LU
// solve for gfx950, reproducing MatFlat's FP64 operation order exactly.
//
// Scheduling (MI355X):
//  * A device-side counting sort orders the batch by size key (largest
//    first) without any host round trip, and splits it into 16-wide size
//    buckets.  Each bucket runs one persistent launch sized to the CUs; its
//    workgroups walk the bucket's items with a fixed stride, so there is no
//    shared work counter (a single L2 address serialises atomics) and
//    neighbouring workgroups hold similar sizes.
//  * The Cholesky / SPD-solve kernel packs several matrices into one
//    wavefront (8, 16, 32 or 64 lanes per matrix by size) with the strict
//    lower triangle packed in LDS, and every lane owns several rows: at
//    step k all rows below k compute their dot product for column k at
//    once, so one wavefront carries 2-8 independent dot-product chains per
//    lane (instruction-level parallelism instead of occupancy, which the
//    LDS footprint limits).
//  * LU, triangular solve and LU solve keep one workgroup per matrix with
//    the matrix in LDS and one thread per row.
//  * Every output element is exactly one Internals.Dot of MatFlat's length
//    and pairing, fed with the same operands (MatFlat/Internals.cs:265-292),
//    so the results are bit-identical to MatFlat's sequential loops.
#include "mfh_device.h"
#include "mfh_internal.h"

#include <climits>

#pragma clang fp contract(off)

namespace {


constexpr int kKeys = MFH_NBUCKETS * MFH_BUCKET_STEP + 1;  // size keys 0..144

__host__ __device__ inline int bucket_of(int key) { return (key + MFH_BUCKET_STEP - 1) / MFH_BUCKET_STEP - 1; }

// ---------------------------------------------------------------- sorting
// Workspace (ints): keys[count] items[count] hist[kKeys] cursor[kKeys]
// bstart[NB] bcount[NB] fstart[NF] fcount[NF].  The coarse buckets (16 keys
// wide) feed the exact kernels, the fine ones (8 keys wide) the FAST kernels,
// whose register tiles come in steps of 8 rows.
constexpr int kFine = (kKeys + MFH_FINE_STEP - 1) / MFH_FINE_STEP;

struct SortWs {
    int* keys;
    int* items;
    int* hist;
    int* cursor;
    int* bstart;
    int* bcount;
    int* fstart;
    int* fcount;
};

size_t sort_ws_bytes(int count)
{
    return sizeof(int) * (2 * (size_t)count + 2 * kKeys + 2 * MFH_NBUCKETS + 2 * kFine);
}

SortWs sort_ws(void* base, int count)
{
    int* p = (int*)base;
    SortWs w;
    w.keys = p;
    w.items = p + count;
    w.hist = p + 2 * (size_t)count;
    w.cursor = w.hist + kKeys;
    w.bstart = w.cursor + kKeys;
    w.bcount = w.bstart + MFH_NBUCKETS;
    w.fstart = w.bcount + MFH_NBUCKETS;
    w.fcount = w.fstart + kFine;
    return w;
}

__global__ void k_sort_reset(SortWs ws)
{
    for (int i = threadIdx.x; i < kKeys; i += blockDim.x) {
        ws.hist[i] = 0;
        ws.cursor[i] = 0;
    }
}

// Validates matrix i through op.key(i) (which writes a negative info and
// returns -1 for an invalid matrix) and counts the key in a block-local
// histogram; one global atomic per key per block.
template <typename Op>
__global__ __launch_bounds__(256) void k_sort_hist(Op op, int count, SortWs ws)
{
    __shared__ int h[kKeys];
    for (int t = threadIdx.x; t < kKeys; t += blockDim.x) h[t] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    if (i < count) {
        key = op.key(i);
        ws.keys[i] = key;
    }
    if (key >= 0) atomicAdd(&h[key], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < kKeys; t += blockDim.x)
        if (h[t]) atomicAdd(&ws.hist[t], h[t]);
}

// Descending-key exclusive offsets, per-key cursors and bucket ranges.  The
// histogram is staged in LDS first: one thread walking it in global memory
// waits for every load in turn.
__global__ __launch_bounds__(64) void k_sort_offsets(SortWs ws)
{
    __shared__ int h[kKeys];
    __shared__ int cur[kKeys];
    for (int k = threadIdx.x; k < kKeys; k += blockDim.x) h[k] = ws.hist[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = kKeys - 1; k >= 0; --k) {
            cur[k] = run;
            run += h[k];
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kKeys; k += blockDim.x) ws.cursor[k] = cur[k];
    for (int b = threadIdx.x; b < MFH_NBUCKETS; b += blockDim.x) {
        const int hi = MFH_BUCKET_STEP * (b + 1), lo = MFH_BUCKET_STEP * b;
        int c = 0;
        for (int k = lo + 1; k <= hi; ++k) c += h[k];
        ws.bstart[b] = cur[hi];
        ws.bcount[b] = c;
    }
    for (int f = threadIdx.x; f < kFine; f += blockDim.x) {
        const int lo = MFH_FINE_STEP * f;
        const int hi = lo + MFH_FINE_STEP < kKeys - 1 ? lo + MFH_FINE_STEP : kKeys - 1;
        int c = 0;
        for (int k = lo + 1; k <= hi; ++k) c += h[k];
        ws.fstart[f] = cur[hi];
        ws.fcount[f] = c;
    }
}

__global__ __launch_bounds__(256) void k_sort_scatter(int count, SortWs ws)
{
    __shared__ int h[kKeys];
    __shared__ int base[kKeys];
    for (int t = threadIdx.x; t < kKeys; t += blockDim.x) h[t] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int key = (i < count) ? ws.keys[i] : -1;
    int lpos = -1;
    if (key >= 0) lpos = atomicAdd(&h[key], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < kKeys; t += blockDim.x)
        if (h[t]) base[t] = atomicAdd(&ws.cursor[t], h[t]);
    __syncthreads();
    if (key >= 0) ws.items[base[key] + lpos] = i;
}

template <typename Op>
int device_sort(mfh_handle* h, const Op& op, int count, SortWs* out)
{
    int st = mfh_ws_reserve(h, sort_ws_bytes(count));
    if (st) return st;
    SortWs ws = sort_ws(h->ws, count);
    const unsigned blocks = (unsigned)((count + 255) / 256);
    st = mfh_timing_begin(h, MFH_KERNEL_SORT, 0);
    if (st) return st;
    hipLaunchKernelGGL(k_sort_reset, dim3(1), dim3(256), 0, h->stream, ws);
    hipLaunchKernelGGL(k_sort_hist<Op>, dim3(blocks), dim3(256), 0, h->stream, op, count, ws);
    hipLaunchKernelGGL(k_sort_offsets, dim3(1), dim3(64), 0, h->stream, ws);
    hipLaunchKernelGGL(k_sort_scatter, dim3(blocks), dim3(256), 0, h->stream, count, ws);
    MFH_HIP_CHECK(hipGetLastError());
    st = mfh_timing_end(h);
    if (st) return st;
    *out = ws;
    return MFH_OK;
}

MfhQueue bucket_queue(const SortWs& ws, int b) { return MfhQueue{ws.bstart + b, ws.bcount + b, ws.items, 0}; }
MfhQueue fine_queue(const SortWs& ws, int f) { return MfhQueue{ws.fstart + f, ws.fcount + f, ws.items, 0}; }

// Launches a persistent kernel for one bucket: as many workgroups as fit
// on the device (by occupancy), capped by what the bucket could need.
template <typename K, typename... Args>
int launch_bucket(mfh_handle* h, int kind, int rows, K kernel, int threads, size_t lds, long max_groups, Args... args)
{
    MFH_HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
    if (per_cu < 1) per_cu = 1;
    long grid = (long)per_cu * h->num_cus;
    if (grid > max_groups) grid = max_groups;
    if (grid < 1) grid = 1;
    int st = mfh_timing_begin(h, kind, rows);
    if (st) return st;
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(threads), lds, h->stream, args...);
    MFH_HIP_CHECK(hipGetLastError());
    return mfh_timing_end(h);
}

int threads_for(int rows) { return MFH_WAVE * ((rows + MFH_WAVE - 1) / MFH_WAVE); }

int nbuckets_for(int max_key)
{
    int nb = bucket_of(max_key) + 1;
    return nb > MFH_NBUCKETS ? MFH_NBUCKETS : nb;
}

int bucket_bound(int b, int max_key)
{
    int hi = MFH_BUCKET_STEP * (b + 1);
    return hi < max_key ? hi : max_key;
}

// ------------------------------------------------------------------ LU
struct GetrfKey {
    MfhGetrfBatch B;
    int max_m, max_n;
    __device__ int key(int i) const
    {
        const int m = B.m ? B.m[i] : B.m0;
        const int n = B.n ? B.n[i] : B.n0;
        const int lda = B.lda ? B.lda[i] : B.lda0;
        const double* a = B.a ? B.a[i] : B.a0;
        const int* piv = B.piv ? B.piv[i] : B.piv0;
        int bad = 0;
        if (m < 1 || m > max_m || m > MFH_EXACT_MAX_N) bad = -3;
        else if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -4;
        else if (a == nullptr) bad = -5;
        else if (lda < m) bad = -6;
        else if (piv == nullptr) bad = -7;
        if (bad) {
            B.info[i] = bad;
            return -1;
        }
        // larger orders run k_getrf_gen (mfh_general.hip)
        return (m > n ? m : n) > MFH_LDS_MAX_N ? -1 : (m > n ? m : n);
    }
};

// Factorization.Lu.cs:296-350 in Crout order: at step k the column k
// candidates (rows i >= k, dots of length k) are computed in parallel, the
// pivot is a (|w|, index) max-reduction, rows are swapped across all n
// columns, the multipliers are divided, and row k of U (columns j > k, dots
// of length k) is computed in parallel.
template <int T>
__global__ __launch_bounds__(T) void k_getrf_exact(MfhGetrfBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    __shared__ double s_rkey[T / MFH_WAVE];
    __shared__ int s_ridx[T / MFH_WAVE];
    __shared__ int s_p;
    const int tid = threadIdx.x;
    for (int it = blockIdx.x;; it += gridDim.x) {
        const int b = mfh_item(q, it);
        if (b < 0) break;
        const int m = B.m ? B.m[b] : B.m0;
        const int n = B.n ? B.n[b] : B.n0;
        const int lda = B.lda ? B.lda[b] : B.lda0;
        double* A = B.a ? B.a[b] : B.a0 + (size_t)b * B.stride_a;
        int* P = B.piv ? B.piv[b] : B.piv0 + (size_t)b * B.stride_piv;
        // m x n in LDS with an odd leading dimension ml: the row-wise
        // accesses (row i across columns, one thread per column) then fall
        // in distinct banks instead of all in one
        const int ml = m | 1;
        double* L = lds;
        int* pv = (int*)(lds + ml * n);

        stage_block<T>(L, ml, A, lda, m, n, tid);
        for (int i = tid; i < m; i += T) pv[i] = i;
        __syncthreads();

        int sign = 1;
        const int kmax = m < n ? m : n;
        for (int k = 0; k < kmax; ++k) {
            const int i = k + tid;
            const bool act = i < m;
            double w = 0.0;
            if (act) {
                w = L[i + k * ml] - mfh_dot_exact(k, L + i, ml, L + k * ml, 1);
                L[i + k * ml] = w;
            }
            double key = act ? mfh_pivot_key(w) : -2.0;
            int p;
            if constexpr (T == MFH_WAVE) {
                // One wavefront: the largest key, then the lowest lane that
                // holds it (lane = row - k, so the lowest row: the strict
                // '>' scan's choice).  Row k is active, so the maximum is a
                // real candidate's key.
                double mk = key;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) mk = __builtin_fmax(mk, __shfl_xor(mk, off, MFH_WAVE));
                const unsigned long long hit = __ballot(key == mk);
                const double wk = __shfl(w, 0, MFH_WAVE);  // lane 0 holds w_k
                p = __builtin_isnan(wk) ? k : k + __ffsll((long long)hit) - 1;
                __syncthreads();
            } else {
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
                p = s_p;
            }
            if (p != k) {
                for (int j = tid; j < n; j += T) {
                    double t = L[k + j * ml];
                    L[k + j * ml] = L[p + j * ml];
                    L[p + j * ml] = t;
                }
                if (tid == 0) {
                    int t = pv[k];
                    pv[k] = pv[p];
                    pv[p] = t;
                }
                sign = -sign;
                __syncthreads();
            }
            const double diag = L[k + k * ml];
            if (act && i > k && diag != 0.0) L[i + k * ml] = L[i + k * ml] / diag;
            const int j = k + 1 + tid;
            if (j < n) L[k + j * ml] = L[k + j * ml] - mfh_dot_exact(k, L + k, ml, L + j * ml, 1);
            __syncthreads();
        }

        for (int e = tid; e < m * n; e += T) {
            int j = e / m;
            int i = e - j * m;
            A[i + (size_t)j * lda] = L[i + j * ml];
        }
        for (int i = tid; i < m; i += T) P[i] = pv[i];
        if (tid == 0) {
            B.sign[b] = sign;
            B.info[b] = 0;
        }
        __syncthreads();
    }
}

// Factorization.Lu.cs:296-350 for m, n <= 64 with two wavefronts per matrix,
// pipelined: at step k wavefront 1 finishes row k-1 of U (U[k-1][j], j >= k,
// the dots the reference forms when it reaches column j) while wavefront 0
// forms the candidates of column k (rows i >= k).  A candidate's dot of
// length k needs U[k-1][k] only in its last pair, (p_{k-2} + p_{k-1}):
// everything before that pair is Internals.Dot of length k-2 (same parity,
// same pairing) and runs concurrently with the row; after the barrier the
// last pair is added, the pivot is found (a DPP wave maximum of the key's
// ordered bits, then a ballot for the lowest row), the column is swapped and
// divided in registers and the other columns' rows k and p are exchanged by
// the same wavefront: two barriers per step.
// Every element is the same sequence of operations as in k_getrf_exact, so
// the factors, pivots and sign are bit-identical; one dot-product latency
// per step instead of two.
__global__ __launch_bounds__(128) void k_getrf_exact_w2(MfhGetrfBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const int lane = tid & (MFH_WAVE - 1);
    const bool w0 = __builtin_amdgcn_readfirstlane(tid >> 6) == 0;
    for (int it = blockIdx.x;; it += gridDim.x) {
        const int b = mfh_item(q, it);
        if (b < 0) break;
        const int m = B.m ? B.m[b] : B.m0;
        const int n = B.n ? B.n[b] : B.n0;
        const int lda = B.lda ? B.lda[b] : B.lda0;
        double* A = B.a ? B.a[b] : B.a0 + (size_t)b * B.stride_a;
        int* P = B.piv ? B.piv[b] : B.piv0 + (size_t)b * B.stride_piv;
        const int ml = m | 1;  // odd leading dimension: row-wise reads in distinct banks
        double* L = lds;
        int* pv = (int*)(lds + ml * n);
        stage_block<128>(L, ml, A, lda, m, n, tid);
        for (int i = tid; i < m; i += 128) pv[i] = i;
        __syncthreads();
        int sign = 1;
        const int kmax = m < n ? m : n;
        for (int k = 0; k < kmax; ++k) {
            const int i = k + lane;
            const bool act = i < m;
            double sp = 0.0, p2 = 0.0, aik = 0.0, xk1 = 0.0;
            if (w0) {
                // all of the candidate's dot but its last pair, and the two
                // entries of the row the rest needs (wavefront 1 writes only
                // row k-1 before the barrier)
                if (act) {
                    aik = L[i + k * ml];
                    if (k >= 1) xk1 = L[i + (k - 1) * ml];
                }
                if (act && k >= 2) {
                    sp = mfh_dot_exact(k - 2, L + i, ml, L + k * ml, 1);
                    p2 = L[i + (k - 2) * ml] * L[(k - 2) + k * ml];
                }
            } else {
                const int j = k + lane;
                if (k >= 1 && j < n)
                    L[(k - 1) + j * ml] = L[(k - 1) + j * ml] - mfh_dot_exact(k - 1, L + (k - 1), ml, L + j * ml, 1);
            }
            __syncthreads();
            if (w0) {
                double w = 0.0;
                if (act) {
                    double s;
                    if (k >= 2) {
                        s = sp + (p2 + xk1 * L[(k - 1) + k * ml]);
                    } else if (k == 1) {
                        s = xk1 * L[ml];  // Internals.Dot of length 1: x0 * y0
                    } else {
                        s = 0.0;
                    }
                    w = aik - s;
                }
                // the largest key, then the lowest lane holding it (the
                // lowest row: the strict '>' scan's choice)
                const unsigned long long key = mfh_pivot_key_u64(act, w);
                const unsigned long long mk = mfh_wave_max_u64(key);
                const unsigned long long hit = __ballot(key == mk);
                const double wk = mfh_readlane_d(w, 0);  // lane 0 holds row k
                const int p = __builtin_isnan(wk) ? k : k + __ffsll((long long)hit) - 1;
                const double wp = mfh_readlane_d(w, p - k);
                // column k after the exchange of rows k and p, then the
                // multipliers (true division, skipped for a zero pivot)
                double v = (i == p) ? wk : ((i == k) ? wp : w);
                if (act && i > k && wp != 0.0) v = v / wp;
                if (act) L[i + k * ml] = v;
                // the exchange of rows k and p in the other columns, one
                // column per lane, while wavefront 1 waits at the barrier
                if (p != k) {
                    for (int j = lane; j < n; j += MFH_WAVE) {
                        if (j == k) continue;
                        const double t = L[k + j * ml];
                        L[k + j * ml] = L[p + j * ml];
                        L[p + j * ml] = t;
                    }
                    if (lane == 0) {
                        const int t = pv[k];
                        pv[k] = pv[p];
                        pv[p] = t;
                    }
                    sign = -sign;  // only thread 0's copy is written out
                }
            }
            __syncthreads();
        }
        // the last row of U (columns beyond the last step, n > m)
        if (!w0) {
            const int j = kmax + lane;
            if (j < n) L[(kmax - 1) + j * ml] = L[(kmax - 1) + j * ml] - mfh_dot_exact(kmax - 1, L + (kmax - 1), ml, L + j * ml, 1);
        }
        __syncthreads();
        for (int e = tid; e < m * n; e += 128) {
            const int j = e / m;
            const int ii = e - j * m;
            A[ii + (size_t)j * lda] = L[ii + j * ml];
        }
        for (int ii = tid; ii < m; ii += 128) P[ii] = pv[ii];
        if (tid == 0) {
            B.sign[b] = sign;
            B.info[b] = 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ Cholesky / posv
struct CholKey {
    MfhCholBatch B;
    int max_n, posv;
    __device__ int key(int i) const
    {
        const int n = B.n[i];
        int bad = 0;
        if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -3;
        else if (B.a[i] == nullptr) bad = -4;
        else if (B.lda[i] < n || (long long)B.lda[i] * n > MFH_MAX_MATRIX_ELEMS) bad = -5;
        else if (posv && B.x[i] == nullptr) bad = -6;
        if (bad) {
            B.info[i] = bad;
            return -1;
        }
        if (n > MFH_LDS_MAX_N) return -1;  // k_potrf_gen (mfh_general.hip)
        return posv ? n + 1 : n;
    }
};

// Packed strict lower triangle of an NB x NB matrix, row-major: element
// (j, c), c < j, at rowp(j) + c.  Row j is contiguous, so the two elements
// of a dot-product pair are adjacent (one ds_read2_b64), and a column of the
// input's upper triangle lands in one contiguous row.
__device__ __forceinline__ int rowp(int j) { return (j * (j - 1)) >> 1; }

// Factorization.Cholesky.cs:114-168 scheduled by columns.  At step k every
// row j > k computes L[j,k] = (A[k,j] - Dot(k, L[j,0:k], L[k,0:k])) / L[k,k]
// and adds t*t to its own running sum (the reference's s, accumulated in
// the same k order); the owner of row k+1 then finishes that diagonal.  The
// strict upper triangle is loaded transposed into the packed strict lower
// triangle, so A[k,j] sits exactly where L[j,k] is written.
//
// Look-ahead: the dot product of step k is Internals.Dot's pairwise sum; all
// of it except the last pair (k-2, k-1) only involves columns finished two
// steps earlier, so each step computes that bulk for the NEXT step while
// the short critical chain of the current step (last pair, subtraction,
// division, the next diagonal's square root) is in flight.  The result is
// the same sum in the same order.
//
// With POSV the right-hand side is carried as an extra row n (b[c] in place
// of A[c,n]): its "L[n,k]" is exactly the forward substitution of
// SolveTriangular(Lower, NoTrans) (Blas.SolveTriangular.cs:216-222; the
// products are the same pairs, multiplied in the other operand order, which
// is exact).  The Lower/Trans sweep (:224-231) then runs bottom-up with one
// lane per unknown.
//
// G lanes serve one matrix (64/G matrices per wavefront); lane l owns rows
// l, l+G, ..., l+(S-1)G.  NB is the packed dimension (rows, including the
// right-hand-side row), per_mat the LDS doubles per matrix.
template <int G, int S, bool POSV>
__global__ __launch_bounds__(64) void k_chol_packed(MfhCholBatch B, MfhQueue q, int NB, int per_mat)
{
    extern __shared__ double lds[];
    constexpr int MW = MFH_WAVE / G;
    const int lane = threadIdx.x;
    const int grp = lane / G;
    const int l = lane - grp * G;
    double* Mx = lds + (size_t)grp * per_mat;
    const int tri = NB * (NB - 1) / 2;
    // Broadcast slots, double-buffered by step parity: a step reads slot
    // [k & 1] while the owner of the next row writes slot [(k + 1) & 1], so
    // no location is read and written by different lanes in the same step.
    double* dgp = Mx + tri;      // diagonal L[k,k] for step k at dgp[k & 1]
    double* xsp = Mx + tri + 2;  // x_j of back-substitution step j at xsp[j & 1]
    int* failp = (int*)(Mx + tri + 4);

    for (int round = 0;; ++round) {
        const int idx0 = (round * (int)gridDim.x + (int)blockIdx.x) * MW;
        if (idx0 >= *q.count) break;
        const int b = mfh_item(q, idx0 + grp);
        const int n = b >= 0 ? B.n[b] : 0;
        const int N = POSV ? n + 1 : n;
        int nmax = n;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            int o = __shfl_xor(nmax, off, MFH_WAVE);
            nmax = o > nmax ? o : nmax;
        }
        // Wave-uniform loop bounds keep every step loop a scalar loop.
        nmax = __builtin_amdgcn_readfirstlane(nmax);
        double* A = b >= 0 ? B.a[b] : nullptr;
        const int lda = b >= 0 ? B.lda[b] : 0;
        double* X = (POSV && b >= 0) ? B.x[b] : nullptr;

        // Column j of the strict upper triangle -> row j (both contiguous):
        // element (c, j), c < j, to Mx[rowp(j) + c].  Lane l of the group
        // walks packed elements l, l + G, ... (row j holds j of them) with
        // eight loads in flight; one load, wait and store per element had
        // serialised n^2 / (2G) HBM round trips per lane.
        {
            typedef const __attribute__((address_space(1))) double* GSrc;
            const int tot = n > 1 ? rowp(n) : 0;
            int j = 1, c = l;
            while (c >= j) {
                c -= j;
                ++j;
            }
            for (int e0 = l; e0 < tot; e0 += 8 * G) {
                double v[8];
                int dst[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool ok = e0 + u * G < tot;
                    dst[u] = ok ? rowp(j) + c : -1;
                    v[u] = ok ? ((GSrc)A)[c + (size_t)j * lda] : 0.0;
                    c += G;
                    while (c >= j && j < n) {
                        c -= j;
                        ++j;
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (dst[u] >= 0) Mx[dst[u]] = v[u];
            }
        }
        if (POSV)
            for (int c = l; c < n; c += G) Mx[rowp(n) + c] = X[c];
        double ajj[S], ss[S], dd[S], lm1[S], lm2[S], bulk[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int j = l + s * G;
            ajj[s] = (j < n) ? A[j + (size_t)j * lda] : 0.0;
            ss[s] = 0.0;
            dd[s] = 0.0;
            lm1[s] = 0.0;
            lm2[s] = 0.0;
            bulk[s] = 0.0;
        }
        if (l == 0) {
            *failp = 0;
            if (n > 0) {
                double d = ajj[0] - 0.0;
                if (d > 0) {
                    dd[0] = __builtin_sqrt(d);
                    dgp[0] = dd[0];
                } else {
                    *failp = 1;
                }
            }
        }
        __syncthreads();
        int fail = (n > 0) ? *(volatile int*)failp : 0;
        const int steps = POSV ? n : n - 1;
        const int smax = POSV ? nmax : nmax - 1;
        for (int k = 0; k < smax; ++k) {
            const bool gact = (k < steps) && (fail == 0);
            double nbulk[S];
#pragma unroll
            for (int s = 0; s < S; ++s) nbulk[s] = bulk[s];
            // (1) look-ahead: Dot(k+1) for rows j > k+1 up to (not including)
            // its last pair (k-1, k): columns 0..k-2, all final.  It is
            // computed first, into nbulk, and only becomes bulk after the
            // column-k finish below has used the previous step's bulk.
            const int L1 = k + 1;
            if (L1 >= 3) {
                const int kr = rowp(L1);
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (s * G + G - 1 <= k + 1) continue;  // slot retired (uniform)
                    int jr = l + s * G;
                    jr = jr < k + 2 ? k + 2 : jr;
                    jr = jr < NB - 1 ? jr : NB - 1;
                    const int pr = rowp(jr);
                    double acc;
                    int c;
                    if (L1 & 1) {
                        acc = Mx[pr] * Mx[kr];
                        c = 1;
                    } else {
                        acc = 0.0;
                        c = 0;
                    }
#pragma unroll 4
                    for (; c <= k - 3; c += 2) {
                        const double p0 = Mx[pr + c] * Mx[kr + c];
                        const double p1 = Mx[pr + c + 1] * Mx[kr + c + 1];
                        acc = acc + (p0 + p1);
                    }
                    nbulk[s] = acc;
                }
            }
            // (2) column k: finish Dot(k) with its last pair and divide.
            const double diag = dgp[k & 1];
            double rk1 = 0.0, rk2 = 0.0;
            if (k >= 2) {
                rk2 = Mx[rowp(k) + k - 2];
                rk1 = Mx[rowp(k) + k - 1];
            } else if (k == 1) {
                rk1 = Mx[0];  // L[1,0]
            }
            if (gact) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int j = l + s * G;
                    if (j > k && j < N) {
                        double dot;
                        if (k == 0) dot = 0.0;
                        else if (k == 1) dot = lm1[s] * rk1;
                        else {
                            const double p0 = lm2[s] * rk2;
                            const double p1 = lm1[s] * rk1;
                            dot = bulk[s] + (p0 + p1);
                        }
                        const double t = (Mx[rowp(j) + k] - dot) / diag;
                        Mx[rowp(j) + k] = t;
                        lm2[s] = lm1[s];
                        lm1[s] = t;
                        if (j < n) {
                            ss[s] = ss[s] + t * t;
                            if (j == k + 1) {
                                const double d = ajj[s] - ss[s];
                                if (d > 0) {
                                    dd[s] = __builtin_sqrt(d);
                                    dgp[(k + 1) & 1] = dd[s];
                                } else {
                                    *failp = j + 1;
                                }
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < S; ++s) bulk[s] = nbulk[s];
            __syncthreads();
            fail = (n > 0) ? *(volatile int*)failp : 0;
        }

        if (b >= 0) {
            if (fail) {
                // MatFlat throws at row fail-1 after writing rows 0..fail-2
                // completely and row fail-1 left of the diagonal.
                const int jf = fail - 1;
                for (int c = 0; c < jf; ++c)
                    for (int i = c + 1 + l; i <= jf; i += G) A[i + (size_t)c * lda] = Mx[rowp(i) + c];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int j = l + s * G;
                    if (j < jf) A[j + (size_t)j * lda] = dd[s];
                }
                if (l == 0) B.info[b] = fail;
            } else {
                for (int c = 0; c < n; ++c) {
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const int i = l + s * G;
                        if (i < n) A[i + (size_t)c * lda] = (i < c) ? 0.0 : (i == c ? dd[s] : Mx[rowp(i) + c]);
                    }
                }
                if (l == 0) B.info[b] = 0;
            }
        }

        if (POSV) {
            const bool ok = (b >= 0) && (fail == 0);
            double xr[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int r = l + s * G;
                xr[s] = (ok && r < n) ? Mx[rowp(n) + r] : 0.0;
                if (ok && r == n - 1) {
                    xr[s] = xr[s] / dd[s];
                    xsp[(n - 1) & 1] = xr[s];
                }
            }
            __syncthreads();
            for (int j = nmax - 1; j >= 1; --j) {
                if (ok && j < n) {
                    const double xj = xsp[j & 1];
                    const int pj = rowp(j);
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const int r = l + s * G;
                        if (r < j) {
                            xr[s] = xr[s] - Mx[pj + r] * xj;
                            if (r == j - 1) {
                                xr[s] = xr[s] / dd[s];
                                xsp[(j - 1) & 1] = xr[s];
                            }
                        }
                    }
                }
                __syncthreads();
            }
            if (ok) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int r = l + s * G;
                    if (r < n) X[r] = xr[s];
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ triangular solve
struct TrsvKey {
    MfhTrsvBatch B;
    int max_n;
    __device__ int key(int i) const
    {
        const int n = B.n[i];
        int bad = 0;
        if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -6;
        else if (B.a[i] == nullptr) bad = -7;
        else if (B.lda[i] < n) bad = -8;
        else if (B.x[i] == nullptr) bad = -9;
        else if (B.incx[i] < 1) bad = -10;
        if (bad) {
            B.info[i] = bad;
            return -1;
        }
        return n > MFH_LDS_MAX_N ? -1 : n;  // larger: k_trsv_gen (mfh_general.hip)
    }
};

// Progressive form of a dot-product substitution: thread i owns unknown i
// and folds the product with x_c into its running Internals.Dot as soon as
// x_c is final, keeping the reference's pairing (length i: odd -> x0*y0 then
// pairs (1,2),(3,4)...; even -> pairs (0,1),(2,3)...).
// c is uniform but the row parity is not, so the three cases (s = p for an
// odd row at c = 0, pend = p, s = s + (pend + p)) are predicates and the
// fold runs without divergent branches.  Rows that are not active (act
// false) keep s and pend.
__device__ __forceinline__ void dot_fold(bool act, int i, int c, double p, double& s, double& pend)
{
    const bool odd = i & 1;
    const bool codd = c & 1;
    const bool add = act && (odd ? (c != 0 && !codd) : codd);
    const bool setp = act && (odd ? codd : !codd);
    const bool sets = act && odd && c == 0;
    const double t = s + (pend + p);
    s = add ? t : (sets ? p : s);
    pend = setp ? p : pend;
}

template <int T>
__global__ __launch_bounds__(T) void k_trsv_exact(MfhTrsvBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const bool upper = B.uplo == MFH_UPPER;
    const bool trans = B.trans == MFH_TRANS;
    const bool unit = B.unit != 0;
    for (int it = blockIdx.x;; it += gridDim.x) {
        const int b = mfh_item(q, it);
        if (b < 0) break;
        const int n = B.n[b];
        const int lda = B.lda[b];
        const int incx = B.incx[b];
        const double* A = B.a[b];
        double* X = B.x[b];
        // the triangle packed by columns: half the LDS of a square, so twice
        // the workgroups per CU in the large buckets
        double* S = lds;
        double* xs = lds + ((n * (n + 1)) >> 1);
        auto at = [upper, n](int i, int j) {
            return upper ? ((j * (j + 1)) >> 1) + i : j * n - ((j * (j - 1)) >> 1) + (i - j);
        };
        if (upper)
            stage_block<T, 1, true>(S, n, A, lda, n, n, tid);
        else
            stage_block<T, 2, true>(S, n, A, lda, n, n, tid);
        double xr = (tid < n) ? X[(size_t)tid * incx] : 0.0;
        __syncthreads();
        if (upper == trans) {
            // Dot forms: Lower/NoTrans (Blas.SolveTriangular.cs:216-222) reads
            // row i of L, Upper/Trans (:201-207) reads column i of U.
            double s = 0.0, pend = 0.0;
            if (tid == 0) {
                xr = xr - 0.0;
                if (!unit) xr = xr / S[at(0, 0)];
                xs[0] = xr;
            }
            __syncthreads();
            for (int c = 0; c < n - 1; ++c) {
                const int i = tid;
                const bool act = i > c && i < n;
                const int ai = act ? i : c;  // inactive rows read a valid slot, unused
                const double aic = upper ? S[at(c, ai)] : S[at(ai, c)];
                dot_fold(act, i, c, aic * xs[c], s, pend);
                if (i == c + 1) {
                    xr = xr - s;
                    if (!unit) xr = xr / S[at(i, i)];
                    xs[i] = xr;
                }
                __syncthreads();
            }
        } else {
            // Axpy forms, bottom-up: Upper/NoTrans (:191-199) subtracts column
            // i of U, Lower/Trans (:224-231) row j of L (MulSub :397-421).
            if (tid == n - 1) {
                if (!unit) xr = xr / S[at(n - 1, n - 1)];
                xs[n - 1] = xr;
            }
            __syncthreads();
            for (int j = n - 1; j >= 1; --j) {
                if (tid < j) {
                    double arj = upper ? S[at(tid, j)] : S[at(j, tid)];
                    xr = xr - arj * xs[j];
                    if (tid == j - 1) {
                        if (!unit) xr = xr / S[at(tid, tid)];
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
struct GetrsKey {
    MfhGetrsBatch B;
    int max_n;
    __device__ int key(int i) const
    {
        const int n = B.n[i];
        int bad = 0;
        if (n < 1 || n > max_n || n > MFH_EXACT_MAX_N) bad = -3;
        else if (B.lu[i] == nullptr) bad = -4;
        else if (B.lda[i] < n) bad = -5;
        else if (B.piv[i] == nullptr) bad = -6;
        else if (B.b[i] == nullptr) bad = -7;
        else if (B.x[i] == nullptr) bad = -8;
        if (bad) {
            B.info[i] = bad;
            return -1;
        }
        return n > MFH_LDS_MAX_N ? -1 : n;  // larger: mfh_general.hip
    }
};

// Factorization.LuInverse.cs:169-192 for one column: permute, unit-lower
// dot-form sweep (the reference divides by the 1.0 it stores on the copy's
// diagonal, which is exact), then the Upper/NoTrans axpy sweep.
template <int T>
__global__ __launch_bounds__(T) void k_getrs_exact(MfhGetrsBatch B, MfhQueue q)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    for (int it = blockIdx.x;; it += gridDim.x) {
        const int b = mfh_item(q, it);
        if (b < 0) break;
        const int n = B.n[b];
        const int lda = B.lda[b];
        const double* LU = B.lu[b];
        const int* P = B.piv[b];
        const int nl = n | 1;  // odd leading dimension: row-wise reads hit distinct banks
        double* S = lds;
        double* xs = lds + nl * n;
        stage_block<T>(S, nl, LU, lda, n, n, tid);
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
            const bool act = i > c && i < n;
            dot_fold(act, i, c, S[(act ? i : c) + c * nl] * xs[c], s, pend);
            if (i == c + 1) {
                xr = xr - s;
                xs[i] = xr;
            }
            __syncthreads();
        }
        if (tid == n - 1) {
            xr = xr / S[(n - 1) * (nl + 1)];
            xs[n - 1] = xr;
        }
        __syncthreads();
        for (int j = n - 1; j >= 1; --j) {
            if (tid < j) {
                xr = xr - S[tid + j * nl] * xs[j];
                if (tid == j - 1) {
                    xr = xr / S[tid + tid * nl];
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

// nb x nb doubles with one spare column (odd leading dimension), nb doubles
// and nb ints
size_t square_lds_bytes(int nb) { return (size_t)nb * nb * 8 + (size_t)nb * 8 + (size_t)nb * 8 + (size_t)nb * 4 + 64; }

// Per-bucket shape of the packed Cholesky kernel: lanes per matrix G and
// rows per lane S (G*S >= the bucket's largest row count).
struct CholShape {
    int G, S;
};

// Built-in shapes, as matflat_amd/autotune.py measured them on MI355X
// (its table is matflat_amd/tuning/gfx950.json).
CholShape chol_shape(int b)
{
    switch (b) {
    case 0: return {8, 2};
    case 1: return {16, 2};
    case 2: return {16, 3};
    case 3: return {32, 2};
    case 8: return {64, 3};
    default: return {64, 2};
    }
}

template <bool POSV>
int run_chol_bucket(mfh_handle* h, int b, int nb_rows, int count, const MfhCholBatch& B, const SortWs& ws)
{
    CholShape sh = chol_shape(b);
    if (h->chol_shape[b][0] > 0 && h->chol_shape[b][0] * h->chol_shape[b][1] >= nb_rows)
        sh = CholShape{h->chol_shape[b][0], h->chol_shape[b][1]};
    int per_mat = nb_rows * (nb_rows - 1) / 2 + 5;
    per_mat = (per_mat + 1) & ~1;
    const int mw = MFH_WAVE / sh.G;
    const size_t lds = (size_t)per_mat * 8 * mw;
    const long max_groups = ((long)count + mw - 1) / mw;
    const MfhQueue q = bucket_queue(ws, b);
#define MFH_CHOL_CASE(G_, S_)                                                                                 \
    if (sh.G == G_ && sh.S == S_)                                                                             \
        return launch_bucket(h, MFH_KERNEL_CHOL_EXACT, nb_rows, k_chol_packed<G_, S_, POSV>, MFH_WAVE, lds, max_groups, B, q, nb_rows, per_mat);
    MFH_CHOL_CASE(8, 2)
    MFH_CHOL_CASE(16, 2)
    MFH_CHOL_CASE(16, 3)
    MFH_CHOL_CASE(32, 2)
    MFH_CHOL_CASE(64, 2)
    MFH_CHOL_CASE(64, 3)
#undef MFH_CHOL_CASE
    return mfh_fail(MFH_ERR_UNSUPPORTED, "no Cholesky kernel for this bucket shape");
}

}  // namespace

bool mfh_chol_shape_ok(int bucket, int g, int s)
{
    const bool built = (g == 8 && s == 2) || (g == 16 && s == 2) || (g == 16 && s == 3) || (g == 32 && s == 2) ||
                       (g == 64 && s == 2) || (g == 64 && s == 3);
    if (!built || bucket < 0 || bucket >= MFH_NBUCKETS) return false;
    // every row of the bucket's largest matrix has a lane, and the packed
    // triangles of the 64 / g matrices of a wavefront fit the LDS
    const long rows = MFH_BUCKET_STEP * (bucket + 1);
    const long per_mat = rows * (rows - 1) / 2 + 6;
    return g * s >= rows && per_mat * 8 * (64 / g) <= 160 * 1024;
}

int mfh_launch_getrf(mfh_handle* h, int count, const MfhGetrfBatch& B, int max_m, int max_n)
{
    const int top = max_m > max_n ? max_m : max_n;
    if (B.a == nullptr && top <= MFH_LDS_MAX_N) {
        // one shape, checked on the host: no sort, one launch over the batch
        const MfhQueue q{nullptr, nullptr, nullptr, count};
        const size_t lds = square_lds_bytes(top);
        if (top <= MFH_WAVE)
            return launch_bucket(h, MFH_KERNEL_GETRF_EXACT, top, k_getrf_exact_w2, 128, lds, count, B, q);
        return launch_bucket(h, MFH_KERNEL_GETRF_EXACT, top, k_getrf_exact<128>, 128, lds, count, B, q);
    }
    SortWs ws;
    int st = device_sort(h, GetrfKey{B, max_m, max_n}, count, &ws);
    if (st) return st;
    const int key = top > MFH_LDS_MAX_N ? MFH_LDS_MAX_N : top;
    for (int b = 0; b < nbuckets_for(key); ++b) {
        const int nb = bucket_bound(b, key);
        const int t = threads_for(nb);
        const size_t lds = square_lds_bytes(nb);
        const MfhQueue q = bucket_queue(ws, b);
        if (t == 64) st = launch_bucket(h, MFH_KERNEL_GETRF_EXACT, nb, k_getrf_exact_w2, 128, lds, count, B, q);
        else st = launch_bucket(h, MFH_KERNEL_GETRF_EXACT, nb, k_getrf_exact<128>, 128, lds, count, B, q);
        if (st) return st;
    }
    if (top > MFH_LDS_MAX_N) return mfh_launch_getrf_gen_d(h, count, B, max_m, max_n);
    return MFH_OK;
}

// Runs launch(i) for the nb size buckets i = nb-1 .. 0.  The buckets are
// independent persistent launches, each ending in a partial last round of
// matrices.  Forked over h->bucket_streams side streams (largest bucket
// first, round robin) the next bucket fills the CUs a finishing one leaves
// idle; the handle's stream waits for all of them.  Under graph capture the
// fork and join become graph edges.  With per-launch timing on, the
// launches run in order on the handle's stream so each is timed alone.
template <typename F>
int fork_buckets(mfh_handle* h, int nb, F launch)
{
    int k = h->timing ? 1 : (h->bucket_streams < nb ? h->bucket_streams : nb);
    // a call that allocates nothing uses only side streams that exist already
    if (h->no_alloc && k > (int)h->side.size()) k = (int)h->side.size();
    if (h->no_alloc && (int)h->side_ev.size() < k + 1) k = 1;
    if (k <= 1) {
        for (int i = nb - 1; i >= 0; --i) {
            int st = launch(i);
            if (st) return st;
        }
        return MFH_OK;
    }
    int st = mfh_side_streams(h, k);
    if (st) return st;
    const hipStream_t main = h->stream;
    MFH_HIP_CHECK(hipEventRecord(h->side_ev[k], main));
    for (int s = 0; s < k; ++s) MFH_HIP_CHECK(hipStreamWaitEvent(h->side[s], h->side_ev[k], 0));
    for (int i = nb - 1, r = 0; i >= 0; --i, ++r) {
        h->stream = h->side[r % k];
        st = launch(i);
        h->stream = main;
        if (st) return st;
    }
    for (int s = 0; s < k; ++s) {
        MFH_HIP_CHECK(hipEventRecord(h->side_ev[s], h->side[s]));
        MFH_HIP_CHECK(hipStreamWaitEvent(main, h->side_ev[s], 0));
    }
    return MFH_OK;
}

int mfh_launch_potrf(mfh_handle* h, int count, const MfhCholBatch& B, int max_n, bool posv)
{
    SortWs ws;
    int st = device_sort(h, CholKey{B, max_n, posv ? 1 : 0}, count, &ws);
    if (st) return st;
    const int n_cap = max_n > MFH_LDS_MAX_N ? MFH_LDS_MAX_N : max_n;
    const int key = posv ? n_cap + 1 : n_cap;
    if (h->math_mode == MFH_MATH_FAST) {
        // fine bucket f holds row counts (8f, 8f + 8]: register tile of f + 1 row blocks
        const int nf = (key + MFH_FINE_STEP - 1) / MFH_FINE_STEP;
        st = fork_buckets(h, nf, [&](int f) {
            return mfh_launch_chol_fast_bucket(h, B, fine_queue(ws, f), f + 1, count, posv);
        });
        if (st) return st;
        // orders above the register tiles run the MatFlat-order kernels
        if (max_n > MFH_LDS_MAX_N) return mfh_launch_potrf_gen_d(h, count, B, max_n, posv);
        return MFH_OK;
    }
    // The exact buckets (16 rows wide) stay in order: one bucket usually
    // dominates, and measured on MI355X the fork only added its waits
    // (n = 32: 0.651 -> 0.674 ms; n ~ U{10..100}: 11.49 -> 11.48 ms).
    for (int b = 0; b < nbuckets_for(key); ++b) {
        const int nb_rows = bucket_bound(b, key) < 2 ? 2 : bucket_bound(b, key);
        st = posv ? run_chol_bucket<true>(h, b, nb_rows, count, B, ws)
                  : run_chol_bucket<false>(h, b, nb_rows, count, B, ws);
        if (st) return st;
    }
    if (max_n > MFH_LDS_MAX_N) return mfh_launch_potrf_gen_d(h, count, B, max_n, posv);
    return MFH_OK;
}

int mfh_launch_trsv(mfh_handle* h, int count, const MfhTrsvBatch& B, int max_n)
{
    SortWs ws;
    int st = device_sort(h, TrsvKey{B, max_n}, count, &ws);
    if (st) return st;
    const int key = max_n > MFH_LDS_MAX_N ? MFH_LDS_MAX_N : max_n;
    for (int b = 0; b < nbuckets_for(key); ++b) {
        const int nb = bucket_bound(b, key);
        const int t = threads_for(nb);
        const size_t lds = ((size_t)nb * (nb + 1) / 2 + nb) * 8 + 64;  // packed triangle + x
        const MfhQueue q = bucket_queue(ws, b);
        if (t == 64) st = launch_bucket(h, MFH_KERNEL_TRSV_EXACT, nb, k_trsv_exact<64>, 64, lds, count, B, q);
        else st = launch_bucket(h, MFH_KERNEL_TRSV_EXACT, nb, k_trsv_exact<128>, 128, lds, count, B, q);
        if (st) return st;
    }
    if (max_n > MFH_LDS_MAX_N) return mfh_launch_trsv_gen_d(h, count, B, max_n);
    return MFH_OK;
}

int mfh_launch_getrs(mfh_handle* h, int count, const MfhGetrsBatch& B, int max_n)
{
    SortWs ws;
    int st = device_sort(h, GetrsKey{B, max_n}, count, &ws);
    if (st) return st;
    const int key = max_n > MFH_LDS_MAX_N ? MFH_LDS_MAX_N : max_n;
    for (int b = 0; b < nbuckets_for(key); ++b) {
        const int nb = bucket_bound(b, key);
        const int t = threads_for(nb);
        const size_t lds = square_lds_bytes(nb);
        const MfhQueue q = bucket_queue(ws, b);
        if (t == 64) st = launch_bucket(h, MFH_KERNEL_GETRS_EXACT, nb, k_getrs_exact<64>, 64, lds, count, B, q);
        else st = launch_bucket(h, MFH_KERNEL_GETRS_EXACT, nb, k_getrs_exact<128>, 128, lds, count, B, q);
        if (st) return st;
    }
    if (max_n > MFH_LDS_MAX_N) return mfh_launch_getrs_gen_d(h, count, B, max_n);
    return MFH_OK;
}

// ------------------------------------------------------------ LU inverse
// Factorization.LuInverse.cs:105-192 sets X = P (X[j, piv[j]] = 1) and runs
// the unit-lower and the upper solve of :156-174 on every column of X, so
// column j of A^-1 is exactly the LU-solve composition of k_getrs_exact
// applied to b = e_j.  Up to n = 64 k_getri_exact (below) runs all columns
// of a matrix in one workgroup.  Larger orders run max_n solves per matrix
// through k_getrs_exact itself: k_getri_setup checks each matrix (the first zero
// diagonal entry of U gives info = j + 1 and leaves the matrix unchanged,
// as the reference throws before writing), copies its factors aside (A is
// overwritten column by column) and writes the solve list; columns j >= n
// and rejected matrices get n = 0, which the solve's own check skips.
struct GetriSolves {
    int* n;
    const double** lu;
    int* lda;
    const int** piv;
    const double** b;
    double** x;
    int* info;
    double* copy;   // count x (max_n x max_n) factors, leading dimension max_n
    double* ident;  // max_n x max_n identity: column j is e_j
};

__global__ __launch_bounds__(256) void k_getri_setup(MfhGetriBatch B, int count, int max_n, GetriSolves S)
{
    __shared__ int s_zero;
    const int i = blockIdx.x, tid = threadIdx.x;
    if (i < max_n)
        for (int k = tid; k < max_n; k += 256) S.ident[k + (size_t)i * max_n] = (k == i) ? 1.0 : 0.0;
    if (i >= count) return;
    const int n = B.n[i];
    const int lda = B.lda[i];
    double* a = B.a[i];
    const int* piv = B.piv[i];
    int bad = 0;
    if (n < 1 || n > max_n) bad = -3;
    else if (a == nullptr) bad = -4;
    else if (lda < n || (long long)lda * n > MFH_MAX_MATRIX_ELEMS) bad = -5;
    else if (piv == nullptr) bad = -6;
    if (tid == 0) s_zero = INT_MAX;
    __syncthreads();
    if (!bad)
        for (int j = tid; j < n; j += 256)
            if (a[j + (size_t)j * lda] == 0.0) atomicMin(&s_zero, j);
    __syncthreads();
    if (!bad && s_zero != INT_MAX) bad = s_zero + 1;
    double* cp = S.copy + (size_t)i * max_n * max_n;
    if (!bad)
        for (int e = tid; e < n * n; e += 256) {
            const int j = e / n;
            const int r = e - j * n;
            cp[r + (size_t)j * max_n] = a[r + (size_t)j * lda];
        }
    for (int j = tid; j < max_n; j += 256) {
        const size_t s = (size_t)i * max_n + j;
        const bool live = !bad && j < n;
        S.n[s] = live ? n : 0;
        S.lu[s] = cp;
        S.lda[s] = max_n;
        S.piv[s] = piv;
        S.b[s] = S.ident + (size_t)j * max_n;
        S.x[s] = live ? a + (size_t)j * lda : nullptr;
    }
    if (tid == 0) B.info[i] = bad;
}

// Factorization.LuInverse.cs:105-192 for n <= 64, one workgroup of eight
// wavefronts per matrix: the factors are staged in LDS once, and all n
// column solves run together.  Lane i owns row i; wavefront q owns columns
// q, q + 8, ... (eight columns per wavefront: 99 VGPRs, and each step's
// work per wavefront is half that of four wavefronts with sixteen columns
// and 182 VGPRs; both fit two workgroups per CU.  1,024 inverses at n = 64:
// 0.384 -> 0.315 ms).  For every (row, column) the arithmetic is exactly that of
// k_getrs_exact with b = e_j: y = P e_j, the unit-lower dot-form sweep
// folded term by term in MatFlat's pairing, then the Upper/NoTrans axpy
// sweep.  Finished entries of X go through LDS (broadcast reads within a
// wavefront); each step ends in one barrier.
constexpr int GETRI_T = 512;
constexpr int GETRI_MAXN = 64;
constexpr int GETRI_COLS = GETRI_T / MFH_WAVE;     // column stride between a wavefront's columns
constexpr int GETRI_JPER = GETRI_MAXN / GETRI_COLS;  // columns per wavefront

// The factors (n x (n|1)) and two rows of X: a step reads only the row of X
// finished in the previous step, and every lane keeps its own rows in
// registers, so X never needs to be whole in LDS.
size_t getri_lds_bytes(int max_n) { return ((size_t)max_n * (max_n | 1) + 2 * (size_t)max_n) * 8; }

__global__ __launch_bounds__(GETRI_T) void k_getri_exact(MfhGetriBatch B, int count, int max_n)
{
    extern __shared__ double lds[];
    __shared__ int s_zero;
    const int tid = threadIdx.x;
    const int i = tid & (MFH_WAVE - 1);
    const int q = tid / MFH_WAVE;
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const int n = B.n[b];
        const int lda = B.lda[b];
        double* a = B.a[b];
        const int* piv = B.piv[b];
        int bad = 0;
        if (n < 1 || n > max_n) bad = -3;
        else if (a == nullptr) bad = -4;
        else if (lda < n || (long long)lda * n > MFH_MAX_MATRIX_ELEMS) bad = -5;
        else if (piv == nullptr) bad = -6;
        if (tid == 0) s_zero = INT_MAX;
        __syncthreads();
        if (!bad)
            for (int j = tid; j < n; j += GETRI_T)
                if (a[j + (size_t)j * lda] == 0.0) atomicMin(&s_zero, j);
        __syncthreads();
        if (!bad && s_zero != INT_MAX) bad = s_zero + 1;
        if (tid == 0) B.info[b] = bad;
        if (bad) {
            __syncthreads();
            continue;
        }
        const int nl = n | 1;
        double* S = lds;
        double* X = lds + n * nl;  // X[p * n + j]: row c of X in slot p = c & 1
        stage_block<GETRI_T>(S, nl, a, lda, n, n, tid);
        const int pr = i < n ? piv[i] : 0;
        const bool pok = pr >= 0 && pr < n;
        double xr[GETRI_JPER], s[GETRI_JPER], pend[GETRI_JPER];
#pragma unroll
        for (int jj = 0; jj < GETRI_JPER; ++jj) {
            const int j = q + GETRI_COLS * jj;
            xr[jj] = pok ? (pr == j ? 1.0 : 0.0) : __builtin_nan("");
            s[jj] = 0.0;
            pend[jj] = 0.0;
        }
        if (i == 0) {
#pragma unroll
            for (int jj = 0; jj < GETRI_JPER; ++jj) {
                const int j = q + GETRI_COLS * jj;
                if (j < n) {
                    xr[jj] = xr[jj] - 0.0;
                    X[j] = xr[jj];  // row 0, slot 0
                }
            }
        }
        __syncthreads();
        const bool odd = i & 1;
        for (int c = 0; c < n - 1; ++c) {
            // dot_fold's cases as selects (c is uniform, the row parity is
            // not): s = p (odd row, c = 0), pend = p, or s = s + (pend + p)
            const bool act = i > c && i < n;
            const bool codd = c & 1;
            const bool add = act && (odd ? (c != 0 && !codd) : codd);
            const bool setp = act && (odd ? codd : !codd);
            const bool sets = act && odd && c == 0;
            const double l = S[act ? i + c * nl : 0];  // rows outside the matrix read a valid slot, unused
            double xc[GETRI_JPER];
#pragma unroll
            for (int jj = 0; jj < GETRI_JPER; ++jj) {
                const int j = q + GETRI_COLS * jj;
                xc[jj] = X[(c & 1) * n + (j < n ? j : 0)];
            }
#pragma unroll
            for (int jj = 0; jj < GETRI_JPER; ++jj) {
                const double p = l * xc[jj];
                const double t = s[jj] + (pend[jj] + p);
                s[jj] = add ? t : (sets ? p : s[jj]);
                pend[jj] = setp ? p : pend[jj];
            }
            if (i == c + 1) {
#pragma unroll
                for (int jj = 0; jj < GETRI_JPER; ++jj) {
                    const int j = q + GETRI_COLS * jj;
                    xr[jj] = xr[jj] - s[jj];
                    if (j < n) X[(i & 1) * n + j] = xr[jj];
                }
            }
            __syncthreads();
        }
        if (i == n - 1) {
            const double d = S[(n - 1) * (nl + 1)];
#pragma unroll
            for (int jj = 0; jj < GETRI_JPER; ++jj) {
                const int j = q + GETRI_COLS * jj;
                if (j < n) {
                    xr[jj] = xr[jj] / d;
                    X[((n - 1) & 1) * n + j] = xr[jj];
                }
            }
        }
        __syncthreads();
        for (int jc = n - 1; jc >= 1; --jc) {
            const bool act = i < jc;
            const double u = S[act ? i + jc * nl : 0];
            double xc[GETRI_JPER];
#pragma unroll
            for (int jj = 0; jj < GETRI_JPER; ++jj) {
                const int j = q + GETRI_COLS * jj;
                xc[jj] = X[(jc & 1) * n + (j < n ? j : 0)];
            }
#pragma unroll
            for (int jj = 0; jj < GETRI_JPER; ++jj) {
                const double t = xr[jj] - u * xc[jj];
                xr[jj] = act ? t : xr[jj];
            }
            if (i == jc - 1) {
                const double d = S[i + i * nl];
#pragma unroll
                for (int jj = 0; jj < GETRI_JPER; ++jj) {
                    const int j = q + GETRI_COLS * jj;
                    xr[jj] = xr[jj] / d;
                    if (j < n) X[(i & 1) * n + j] = xr[jj];
                }
            }
            __syncthreads();
        }
        if (i < n) {
#pragma unroll
            for (int jj = 0; jj < GETRI_JPER; ++jj) {
                const int j = q + GETRI_COLS * jj;
                if (j < n) a[i + (size_t)j * lda] = xr[jj];
            }
        }
        __syncthreads();
    }
}

// Workspace bytes a batch of count matrices needs (plans reserve this).
size_t mfh_exact_ws_bytes(int count) { return sort_ws_bytes(count); }

int mfh_launch_getri(mfh_handle* h, int count, const MfhGetriBatch& B, int max_n)
{
    if (max_n <= GETRI_MAXN) {
        const size_t lds = getri_lds_bytes(max_n);
        MFH_HIP_CHECK(hipFuncSetAttribute((const void*)k_getri_exact, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds));
        int per_cu = 0;
        MFH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_getri_exact, GETRI_T, lds));
        if (per_cu < 1) per_cu = 1;
        long grid = (long)per_cu * h->num_cus;
        if (grid > count) grid = count;
        int st = mfh_timing_begin(h, MFH_KERNEL_GETRI_EXACT, max_n);
        if (st) return st;
        hipLaunchKernelGGL(k_getri_exact, dim3((unsigned)grid), dim3(GETRI_T), lds, h->stream, B, count, max_n);
        MFH_HIP_CHECK(hipGetLastError());
        return mfh_timing_end(h);
    }
    const size_t solves = (size_t)count * max_n;
    size_t off = 0;
    auto take = [&off](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_copy = take(solves * max_n * sizeof(double));
    const size_t o_ident = take((size_t)max_n * max_n * sizeof(double));
    const size_t o_n = take(solves * sizeof(int)), o_lda = take(solves * sizeof(int));
    const size_t o_info = take(solves * sizeof(int));
    const size_t o_lu = take(solves * sizeof(void*)), o_piv = take(solves * sizeof(void*));
    const size_t o_b = take(solves * sizeof(void*)), o_x = take(solves * sizeof(void*));
    int st = mfh_aux_reserve(h, off);
    if (st) return st;
    char* base = (char*)h->aux;
    GetriSolves S{(int*)(base + o_n),       (const double**)(base + o_lu), (int*)(base + o_lda),
                  (const int**)(base + o_piv), (const double**)(base + o_b), (double**)(base + o_x),
                  (int*)(base + o_info),    (double*)(base + o_copy),      (double*)(base + o_ident)};
    st = mfh_timing_begin(h, MFH_KERNEL_GETRI_SETUP, max_n);
    if (st) return st;
    const unsigned grid = (unsigned)(count > max_n ? count : max_n);
    hipLaunchKernelGGL(k_getri_setup, dim3(grid), dim3(256), 0, h->stream, B, count, max_n, S);
    MFH_HIP_CHECK(hipGetLastError());
    st = mfh_timing_end(h);
    if (st) return st;
    MfhGetrsBatch G{S.n, S.lu, S.lda, S.piv, S.b, S.x, S.info};
    return mfh_launch_getrs(h, (int)solves, G, max_n);
}
