// mfh_order.h -- largest-first order of a batch by the matrices' orders n.
//
// A counting sort of the matrix ids by n / 8, descending, in three small
// kernels (histogram, offsets, scatter) with LDS-aggregated atomics; no host
// round trip, and nothing in it is a memset, so a graph can capture it.
// Orders outside 1 .. MAXN get key 0 and sort last: the kernels that walk
// the order report them.  Ids of one key keep no particular order (the
// matrices are independent, so results do not depend on it).  A persistent
// kernel that takes the ids through an atomic work counter then starts the
// largest matrices first and ends on the small ones, so its workgroups
// finish together instead of some holding several large matrices at the end.
#pragma once

#include "mfh_device.h"

namespace {

template <int MAXN>
struct MfhOrder {
    static constexpr int keys = MAXN / 8 + 1;  // ints of the histogram and of the cursors
    __device__ __forceinline__ static int key(const int* n, int i)
    {
        const int v = n[i];
        return (v >= 1 && v <= MAXN) ? v >> 3 : 0;
    }
};

template <int MAXN>
__global__ __launch_bounds__(256) void k_order_hist(const int* n, int count, int* hist)
{
    constexpr int K = MfhOrder<MAXN>::keys;
    __shared__ int hl[K];
    for (int k = threadIdx.x; k < K; k += blockDim.x) hl[k] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) atomicAdd(&hl[MfhOrder<MAXN>::key(n, i)], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x)
        if (hl[k]) atomicAdd(&hist[k], hl[k]);
}

// cursor[k] = first position of key k: the keys above it come first.
template <int MAXN>
__global__ void k_order_offsets(const int* hist, int* cursor)
{
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = MfhOrder<MAXN>::keys - 1; k >= 0; --k) {
            cursor[k] = run;
            run += hist[k];
        }
    }
}

template <int MAXN>
__global__ __launch_bounds__(256) void k_order_scatter(const int* n, int count, int* cursor, int* order)
{
    constexpr int K = MfhOrder<MAXN>::keys;
    __shared__ int hl[K];
    __shared__ int base[K];
    for (int k = threadIdx.x; k < K; k += blockDim.x) hl[k] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int key = i < count ? MfhOrder<MAXN>::key(n, i) : -1;
    int pos = -1;
    if (key >= 0) pos = atomicAdd(&hl[key], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x)
        if (hl[k]) base[k] = atomicAdd(&cursor[k], hl[k]);
    __syncthreads();
    if (key >= 0) order[base[key] + pos] = i;
}

// Zeroes c[0 .. n) (a kernel, not a memset: the launch sequence may be
// captured into a graph).
__global__ void k_zero_ints(int* c, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = 0;
}

// Writes order[0 .. count) on stream s.  hist and cursor (MfhOrder<MAXN>::keys
// ints each) must be zero on entry.
template <int MAXN>
inline void mfh_launch_order(hipStream_t s, const int* n, int count, int* hist, int* cursor, int* order)
{
    const unsigned blocks = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL(k_order_hist<MAXN>, dim3(blocks), dim3(256), 0, s, n, count, hist);
    hipLaunchKernelGGL(k_order_offsets<MAXN>, dim3(1), dim3(64), 0, s, hist, cursor);
    hipLaunchKernelGGL(k_order_scatter<MAXN>, dim3(blocks), dim3(256), 0, s, n, count, cursor, order);
}

}  // namespace
