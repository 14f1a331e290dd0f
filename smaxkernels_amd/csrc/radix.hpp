// Stable LSD radix sort of (uint32 key, int32 value) pairs on the GPU, 4-bit
// digits.  Each pass: per-tile digit histograms, one prefix sum over the
// digit-major histogram table, then a scatter whose ranks inside a tile come
// from prefix sums only, so equal keys keep their input order and the
// result is deterministic.
#pragma once

#include <utility>

#include "common.hpp"
#include "scan.hpp"

namespace smaxhip {
namespace {  // kernels get internal linkage in every including unit

constexpr int kRThreads = 256;
constexpr int kRItems = 16;
constexpr int kRTile = kRThreads * kRItems;
constexpr int kRBuckets = 16;

__global__ __launch_bounds__(kRThreads) void radix_hist(
    int64_t n, int shift, const uint32_t *__restrict__ keys,
    int32_t *__restrict__ hist, int64_t n_tiles) {
    __shared__ int32_t c[kRBuckets];
    if (threadIdx.x < kRBuckets) c[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kRTile +
                         static_cast<int64_t>(threadIdx.x) * kRItems;
    for (int k = 0; k < kRItems; ++k) {
        const int64_t i = base + k;
        if (i < n) atomicAdd(&c[(keys[i] >> shift) & (kRBuckets - 1)], 1);
    }
    __syncthreads();
    if (threadIdx.x < kRBuckets)
        hist[threadIdx.x * n_tiles + blockIdx.x] = c[threadIdx.x];
}

// incl: inclusive scan of hist (digit-major), so (digit d, tile t) starts at
// incl[d*n_tiles+t] - hist[d*n_tiles+t].
__global__ __launch_bounds__(kRThreads) void radix_scatter(
    int64_t n, int shift, const uint32_t *__restrict__ keys,
    const int32_t *__restrict__ vals, const int32_t *__restrict__ incl,
    const int32_t *__restrict__ hist, int64_t n_tiles,
    uint32_t *__restrict__ okeys, int32_t *__restrict__ ovals) {
    __shared__ int32_t cnt[kRBuckets * kRThreads];
    const int tid = threadIdx.x;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kRTile +
                         static_cast<int64_t>(tid) * kRItems;
    uint32_t kk[kRItems];
    int32_t vv[kRItems];
    int32_t mine[kRBuckets];
    for (int d = 0; d < kRBuckets; ++d) mine[d] = 0;
    for (int k = 0; k < kRItems; ++k) {
        const int64_t i = base + k;
        if (i < n) {
            kk[k] = keys[i];
            vv[k] = vals[i];
            mine[(kk[k] >> shift) & (kRBuckets - 1)]++;
        }
    }
    for (int d = 0; d < kRBuckets; ++d) cnt[d * kRThreads + tid] = mine[d];
    __syncthreads();
    // flat exclusive scan of cnt (digit-major): thread t owns [16t, 16t+16)
    int32_t loc[kRItems];
    int32_t s = 0;
    for (int k = 0; k < kRItems; ++k) {
        loc[k] = s;
        s += cnt[tid * kRItems + k];
    }
    int32_t total;
    const int32_t before = block_excl_scan(s, &total);
    __syncthreads();
    for (int k = 0; k < kRItems; ++k) cnt[tid * kRItems + k] = before + loc[k];
    __syncthreads();
    int32_t run[kRBuckets];
    for (int d = 0; d < kRBuckets; ++d)
        run[d] = incl[d * n_tiles + blockIdx.x] - hist[d * n_tiles + blockIdx.x] +
                 cnt[d * kRThreads + tid] - cnt[d * kRThreads];
    for (int k = 0; k < kRItems; ++k) {
        const int64_t i = base + k;
        if (i < n) {
            const int d = (kk[k] >> shift) & (kRBuckets - 1);
            const int32_t p = run[d]++;
            okeys[p] = kk[k];
            ovals[p] = vv[k];
        }
    }
}

__global__ void iota_keys(int64_t n, const int32_t *__restrict__ col,
                          uint32_t *__restrict__ keys, int32_t *__restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = static_cast<uint32_t>(col[i]);
    vals[i] = static_cast<int32_t>(i);
}


// t_row_ptr[c] = first sorted position with key >= c, for c in [0, n_cols].
__global__ void transpose_row_ptr(int64_t n_cols, int64_t nnz,
                                  const uint32_t *__restrict__ keys,
                                  int32_t *__restrict__ t_row_ptr) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c > n_cols) return;
    int64_t a = 0, b = nnz;
    while (a < b) {
        const int64_t m = (a + b) >> 1;
        if (keys[m] < static_cast<uint32_t>(c)) a = m + 1; else b = m;
    }
    t_row_ptr[c] = static_cast<int32_t>(a);
}

__global__ void transpose_cols(int64_t nnz, int64_t n_rows,
                               const int32_t *__restrict__ row_ptr,
                               const int32_t *__restrict__ src,
                               int32_t *__restrict__ t_col) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int64_t j = src[p];
    int64_t lo = 0, hi = n_rows;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (row_ptr[mid] <= j) lo = mid; else hi = mid;
    }
    t_col[p] = static_cast<int32_t>(lo);
}

// Owner of the ping-pong buffers; keys()/vals() hold the current order.
class RadixSort {
    smax_hip_ctx ctx_;
    uint32_t *k_[2] = {nullptr, nullptr};
    int32_t *v_[2] = {nullptr, nullptr};
    int32_t *hist_ = nullptr, *incl_ = nullptr, *tmp_ = nullptr;
    int cur_ = 0;
    int64_t cap_ = 0;

  public:
    explicit RadixSort(smax_hip_ctx ctx) : ctx_(ctx) {}
    ~RadixSort() {
        (void)hipStreamSynchronize(ctx_->stream);
        for (int i = 0; i < 2; ++i) {
            dev_free(k_[i]);
            dev_free(v_[i]);
        }
        dev_free(hist_);
        dev_free(incl_);
        dev_free(tmp_);
    }
    uint32_t *keys() { return k_[cur_]; }
    int32_t *vals() { return v_[cur_]; }

    int reserve(int64_t n) {
        cap_ = n;
        const int64_t tiles = ceil_div(n, kRTile);
        int rc = 0;
        for (int i = 0; i < 2 && !rc; ++i) {
            rc = dev_alloc(ctx_, reinterpret_cast<void **>(&k_[i]), 4 * n, "radix sort");
            if (!rc) rc = dev_alloc(ctx_, reinterpret_cast<void **>(&v_[i]), 4 * n, "radix sort");
        }
        if (!rc)
            rc = dev_alloc(ctx_, reinterpret_cast<void **>(&hist_), 4 * kRBuckets * tiles,
                           "radix sort");
        if (!rc)
            rc = dev_alloc(ctx_, reinterpret_cast<void **>(&incl_), 4 * kRBuckets * tiles,
                           "radix sort");
        if (!rc)
            rc = dev_alloc(ctx_, reinterpret_cast<void **>(&tmp_),
                           4 * scan_tmp_elems(kRBuckets * tiles), "radix sort");
        return rc;
    }

    // Sorts the first n pairs of keys()/vals() by key; keys < key_bound.
    int sort(int64_t n, int64_t key_bound) {
        if (n <= 0) return SMAX_HIP_SUCCESS;
        const int64_t tiles = ceil_div(n, kRTile);
        int bits = 1;
        while (bits < 32 && (1ll << bits) < key_bound) ++bits;
        for (int shift = 0; shift < bits; shift += 4) {
            hipLaunchKernelGGL(radix_hist, dim3(static_cast<unsigned>(tiles)),
                               dim3(kRThreads), 0, ctx_->stream, n, shift, k_[cur_],
                               hist_, tiles);
            int rc = check_launch(ctx_, "radix_hist");
            if (!rc)
                rc = scan_inclusive<int32_t, false>(ctx_, hist_, incl_,
                                                    kRBuckets * tiles, tmp_);
            if (rc) return rc;
            hipLaunchKernelGGL(radix_scatter, dim3(static_cast<unsigned>(tiles)),
                               dim3(kRThreads), 0, ctx_->stream, n, shift, k_[cur_],
                               v_[cur_], incl_, hist_, tiles, k_[1 - cur_],
                               v_[1 - cur_]);
            rc = check_launch(ctx_, "radix_scatter");
            if (rc) return rc;
            cur_ = 1 - cur_;
        }
        return SMAX_HIP_SUCCESS;
    }
};

}  // namespace
}  // namespace smaxhip
