// Deterministic inclusive prefix sums over int32 / int64 device arrays.
//
// Three passes over 2048-element tiles: per-tile totals, an exclusive scan
// of the totals by one workgroup, then a tile-local scan plus offset.  The
// reverse variant computes suffix sums (out[i] = in[i] + ... + in[n-1]).
// In-place use (in == out) is allowed.
#pragma once

#include "common.hpp"

namespace smaxhip {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

inline int64_t scan_tmp_elems(int64_t n) { return ceil_div(n, kScanTile) + 1; }

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T o = __shfl_up(v, d, kWave);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive scan of one value per thread across the workgroup; returns the
// exclusive prefix and writes the workgroup total to *total.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *total) {
    __shared__ T wsum[kScanThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    T inc = wave_incl_scan(v);
    if (lane == kWave - 1) wsum[wave] = inc;
    __syncthreads();
    T before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / kWave; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

template <typename T, bool REV>
__device__ __forceinline__ int64_t scan_pos(int64_t i, int64_t n) {
    return REV ? n - 1 - i : i;
}

template <typename T, bool REV>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_tiles(
    const T *__restrict__ in, T *__restrict__ part, int64_t n) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + static_cast<int64_t>(k) * kScanThreads + threadIdx.x;
        if (i < n) s += in[scan_pos<T, REV>(i, n)];
    }
    T total;
    (void)block_excl_scan(s, &total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void scan_partials(T *part,
                                                              int64_t m) {
    T carry = 0;
    for (int64_t b = 0; b < m; b += kScanThreads) {
        int64_t i = b + threadIdx.x;
        T v = i < m ? part[i] : T(0);
        T total;
        T ex = block_excl_scan(v, &total);
        if (i < m) part[i] = carry + ex;
        carry += total;
    }
}

template <typename T, bool REV>
__global__ __launch_bounds__(kScanThreads) void scan_tiles(
    const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ part,
    int64_t n) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile +
                         static_cast<int64_t>(threadIdx.x) * kScanItems;
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + k;
        v[k] = i < n ? in[scan_pos<T, REV>(i, n)] : T(0);
        s += v[k];
    }
    T total;
    T run = block_excl_scan(s, &total) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + k;
        run += v[k];
        if (i < n) out[scan_pos<T, REV>(i, n)] = run;
    }
}

// tmp must hold scan_tmp_elems(n) elements of T.
template <typename T, bool REV>
int scan_inclusive(smax_hip_ctx ctx, const T *in, T *out, int64_t n, T *tmp) {
    if (n <= 0) return SMAX_HIP_SUCCESS;
    const int64_t tiles = ceil_div(n, kScanTile);
    hipLaunchKernelGGL((scan_reduce_tiles<T, REV>),
                       dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads),
                       0, ctx->stream, in, tmp, n);
    SMAX_HIP_LAUNCHED(ctx, "scan_reduce_tiles");
    hipLaunchKernelGGL((scan_partials<T>), dim3(1), dim3(kScanThreads), 0,
                       ctx->stream, tmp, tiles);
    SMAX_HIP_LAUNCHED(ctx, "scan_partials");
    hipLaunchKernelGGL((scan_tiles<T, REV>), dim3(static_cast<unsigned>(tiles)),
                       dim3(kScanThreads), 0, ctx->stream, in, out, tmp, n);
    SMAX_HIP_LAUNCHED(ctx, "scan_tiles");
    return SMAX_HIP_SUCCESS;
}

}  // namespace smaxhip
