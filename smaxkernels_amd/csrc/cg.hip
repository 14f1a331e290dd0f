// Conjugate gradients on a SELL-C-sigma matrix: plain FP64 CG and mixed-
// precision iterative refinement (FP64 outer residual / update / norms, inner
// CG on FP32- or BF16-stored values with FP32 vectors and accumulation).
//
// Vectors live in "slot space" (length n_rows_padded, SELL order): the
// solver uses the symmetric permutation P A P^T, whose column indices
// col_slot[e] = perm[col[e]] are built once; every SpMV then writes its
// output coalesced and all vector kernels are contiguous streams.  Padding
// slots have empty rows and stay zero.  b and x are gathered/scattered
// through perm at entry and exit.
//
// BF16 storage rounds the diagonal of the 27-point operator (26.01 -> 26.0)
// and stalls refinement; with keep_diag_f32 the diagonal is taken out of the
// BF16 array (stored as zero) and applied from an FP32 side array as
// y_s += d_s x_s after the row's chain.
//
// Iterations run in batches without host round trips: alpha, beta and the
// dot products stay on the device, the reduction that produces r.r raises a
// device flag once sqrt(r.r) <= stop, and every later kernel of the batch
// returns at once.  Every dot product is a fixed two-pass reduction
// (per-block tree, then one block summing the partials in index order), so a
// solve is deterministic for a given matrix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "convert_internal.hpp"

namespace smaxhip {

constexpr int kCgThreads = 256;
constexpr int64_t kUpdBlocks = 4096;  // fixed grid of the vector kernels

// Device control words of one CG loop.
// CTL_BREAK: the last stop came from a p.q <= 0 breakdown.
enum : int { CTL_DONE = 0, CTL_ITERS = 1, CTL_BREAK = 2 };

template <typename W>
__device__ __forceinline__ W block_sum(W v) {
    __shared__ W ws[kCgThreads / kWave];
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    const int wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) ws[wave] = v;
    __syncthreads();
    W t = ws[0];
#pragma unroll
    for (int w = 1; w < kCgThreads / kWave; ++w) t += ws[w];
    __syncthreads();
    return t;
}

// Sum over a workgroup of NT threads (fixed tree: wave butterflies, then the
// wave sums in wave order).
template <typename W, int NT>
__device__ __forceinline__ W block_sum_n(W v) {
    __shared__ W ws[NT / kWave];
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    const int wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) ws[wave] = v;
    __syncthreads();
    W t = ws[0];
#pragma unroll
    for (int w = 1; w < NT / kWave; ++w) t += ws[w];
    __syncthreads();
    return t;
}

// Reductions of per-block partials run in one workgroup of kReduceThreads;
// each thread issues kReduceUnroll loads before adding them (independent
// accumulators summed in a fixed order), so the pass is not a chain of
// dependent memory round trips.
constexpr int kReduceThreads = 1024;
constexpr int kReduceUnroll = 8;

template <typename W>
__device__ __forceinline__ W reduce_partials(const W *__restrict__ partial, int64_t m) {
    W a[kReduceUnroll];
#pragma unroll
    for (int u = 0; u < kReduceUnroll; ++u) a[u] = 0;
    int64_t i = threadIdx.x;
    constexpr int64_t kStep = static_cast<int64_t>(kReduceThreads) * kReduceUnroll;
    for (; i + (kReduceUnroll - 1) * kReduceThreads < m; i += kStep) {
        W v[kReduceUnroll];
#pragma unroll
        for (int u = 0; u < kReduceUnroll; ++u) v[u] = partial[i + u * kReduceThreads];
#pragma unroll
        for (int u = 0; u < kReduceUnroll; ++u) a[u] += v[u];
    }
#pragma unroll
    for (int u = 0; u < kReduceUnroll; ++u)
        if (i + u * kReduceThreads < m) a[u] += partial[i + u * kReduceThreads];
    W t = a[0];
#pragma unroll
    for (int u = 1; u < kReduceUnroll; ++u) t += a[u];
    return block_sum_n<W, kReduceThreads>(t);
}

template <typename V> struct widen_t;
template <> struct widen_t<double> {
    __device__ static double f(double v) { return v; }
};
template <> struct widen_t<float> {
    __device__ static float f(float v) { return v; }
};
template <> struct widen_t<uint16_t> {
    __device__ static float f(uint16_t v) { return bf16_to_f32(v); }
};

__device__ __forceinline__ double fma_w(double a, double b, double c) {
    return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fma_w(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
}

// Column codings of the solver matrix (every one yields the same column for
// every element, so the products are identical bit for bit):
//   kCol32:  32-bit slot-space columns;
//   kRel16:  16-bit offsets from the slot (rel[e] = col_slot[e] - s; padding
//            elements hold 0 and read the slot's own x entry, times a zero
//            value);
//   kTab8:   C = 64 only: a one-byte index per element into its chunk's table
//            of kColTable 16-bit offsets; lane t of the chunk's wavefront
//            holds entry t mod kColTable and a lane shuffle fetches the entry.
//            A slot's indices of elements j..j+3 share one 32-bit word, so a
//            wavefront reads the indices of four elements per 256-byte load.
enum { kCol32 = 0, kRel16 = 1, kTab8 = 2 };
constexpr int kColTable = 32;

// (A x)[s]: sequential fma chain over the slot's SELL elements, then (DIAG)
// the FP32 diagonal term.  With kTab8 every lane of the wavefront must run
// this (one chunk per wavefront, all slots present).
template <typename V, typename W, bool DIAG, int MODE>
__device__ __forceinline__ W apply_slot(int64_t s, int64_t C,
                                        const int32_t *__restrict__ chunk_ptr,
                                        const int32_t *__restrict__ chunk_len,
                                        const int32_t *__restrict__ col,
                                        const int16_t *__restrict__ rel,
                                        const uint32_t *__restrict__ idx,
                                        const int64_t *__restrict__ idx_row,
                                        const int16_t *__restrict__ tab,
                                        const V *__restrict__ val,
                                        const float *__restrict__ diag,
                                        const W *__restrict__ x) {
    const int64_t c = s / C;
    const int64_t base = static_cast<int64_t>(chunk_ptr[c]) + (s - c * C);
    const int len = chunk_len[c];
    int entry = 0;
    const uint32_t *iw = nullptr;  // kTab8: this slot's index words, one per row
    if constexpr (MODE == kTab8) {
        entry = tab[c * kColTable + (threadIdx.x & (kColTable - 1))];
        iw = idx + idx_row[c] * kWave + (s - c * C);
    }
    // kTab8: the column of element j, its index byte taken from word w
    auto tab_col = [&](uint32_t w, int j) -> int64_t {
        return s + __shfl(entry, static_cast<int>((w >> (8 * (j & 3))) & 0xffu));
    };
    auto col_at = [&](int64_t o) -> int64_t {
        if constexpr (MODE == kRel16)
            return s + ld_stream(rel + o);
        else
            return ld_stream(col + o);
    };
    W acc = 0;
    constexpr int U = 8;
    int j = 0;
    for (; j + U <= len; j += U) {
        int64_t cc[U];
        V vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) vv[u] = ld_stream(val + base + (j + u) * C);
        if constexpr (MODE == kTab8) {
            // j is a multiple of 8: two whole index words
            uint32_t w[U / 4];
#pragma unroll
            for (int g = 0; g < U / 4; ++g) w[g] = ld_stream(iw + ((j >> 2) + g) * kWave);
#pragma unroll
            for (int u = 0; u < U; ++u) cc[u] = tab_col(w[u >> 2], u);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) cc[u] = col_at(base + (j + u) * C);
        }
        W xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[cc[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            acc = fma_w(static_cast<W>(widen_t<V>::f(vv[u])), xv[u], acc);
    }
    for (; j < len; ++j) {
        int64_t cj;
        if constexpr (MODE == kTab8)
            cj = tab_col(ld_stream(iw + (j >> 2) * kWave), j);
        else
            cj = col_at(base + j * C);
        acc = fma_w(static_cast<W>(widen_t<V>::f(val[base + j * C])), x[cj], acc);
    }
    if (DIAG) acc = fma_w(static_cast<W>(diag[s]), x[s], acc);
    return acc;
}

// q = A p, partial[block] = sum p.q
template <typename V, typename W, bool DIAG, int MODE>
__global__ __launch_bounds__(kCgThreads) void k_spmv_dot(
    int64_t n, int64_t C, const int32_t *__restrict__ chunk_ptr,
    const int32_t *__restrict__ chunk_len, const int32_t *__restrict__ col,
    const int16_t *__restrict__ rel, const uint32_t *__restrict__ idx,
    const int64_t *__restrict__ idx_row, const int16_t *__restrict__ tab,
    const V *__restrict__ val,
    const float *__restrict__ diag, const W *__restrict__ p, W *__restrict__ q,
    W *__restrict__ partial, const int64_t *__restrict__ ctl) {
    if (ctl[CTL_DONE]) return;
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
    W prod = 0;
    if (s < n) {
        const W v = apply_slot<V, W, DIAG, MODE>(s, C, chunk_ptr, chunk_len, col, rel, idx,
                                                 idx_row, tab, val, diag, p);
        q[s] = v;
        prod = p[s] * v;
    }
    const W t = block_sum(prod);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// r = b - A x, partial = sum r.r  (FP64 outer residual)
template <int MODE>
__global__ __launch_bounds__(kCgThreads) void k_residual(
    int64_t n, int64_t C, const int32_t *__restrict__ chunk_ptr,
    const int32_t *__restrict__ chunk_len, const int32_t *__restrict__ col,
    const int16_t *__restrict__ rel, const uint32_t *__restrict__ idx,
    const int64_t *__restrict__ idx_row, const int16_t *__restrict__ tab,
    const double *__restrict__ val,
    const double *__restrict__ b, const double *__restrict__ x, double *__restrict__ r,
    double *__restrict__ partial) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
    double sq = 0;
    if (s < n) {
        const double ax = apply_slot<double, double, false, MODE>(
            s, C, chunk_ptr, chunk_len, col, rel, idx, idx_row, tab, val, nullptr, x);
        const double v = b[s] - ax;
        r[s] = v;
        sq = v * v;
    }
    const double t = block_sum(sq);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// *out = sum of partial[0..m) in a fixed order.  With ctl, skipped once the
// loop is done, and (stop2 >= 0) raises the done flag when *out <= stop2.
template <typename W>
__global__ __launch_bounds__(kReduceThreads) void k_reduce(const W *__restrict__ partial,
                                                           int64_t m, W *__restrict__ out,
                                                           int64_t *__restrict__ ctl,
                                                           W stop2) {
    if (ctl && ctl[CTL_DONE]) return;
    const W t = reduce_partials(partial, m);
    if (threadIdx.x == 0) {
        *out = t;
        if (ctl && stop2 >= W(0) && !(t > stop2)) ctl[CTL_DONE] = 1;
    }
}

// *out = p.q (as k_reduce), and the loop stops before alpha is used when p.q
// is not positive (breakdown: the operator is not positive definite in the
// working precision, or a NaN).
template <typename W>
__global__ __launch_bounds__(kReduceThreads) void k_reduce_pq(const W *__restrict__ partial,
                                                              int64_t m, W *__restrict__ out,
                                                              int64_t *__restrict__ ctl) {
    if (ctl[CTL_DONE]) return;
    const W t = reduce_partials(partial, m);
    if (threadIdx.x == 0) {
        *out = t;
        if (!(t > W(0))) {
            ctl[CTL_DONE] = 1;
            ctl[CTL_BREAK] = 1;
        }
    }
}

// x += alpha p, r -= alpha q, partial = sum r.r; alpha = scal[rr] / scal[pq]
template <typename W>
__global__ __launch_bounds__(kCgThreads) void k_update_xr(
    int64_t n, const W *__restrict__ scal, int rr, int pq, W *__restrict__ x,
    W *__restrict__ r, const W *__restrict__ p, const W *__restrict__ q,
    W *__restrict__ partial, int64_t *__restrict__ ctl) {
    if (ctl[CTL_DONE]) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[CTL_ITERS] += 1;
    const W alpha = scal[rr] / scal[pq];
    W sq = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kCgThreads) {
        x[i] = fma_w(alpha, p[i], x[i]);
        const W v = fma_w(-alpha, q[i], r[i]);
        r[i] = v;
        sq = fma_w(v, v, sq);
    }
    const W t = block_sum(sq);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// p = r + beta p, beta = scal[rr_new] / scal[rr_old]
template <typename W>
__global__ __launch_bounds__(kCgThreads) void k_update_p(
    int64_t n, const W *__restrict__ scal, int rr_new, int rr_old,
    const W *__restrict__ r, W *__restrict__ p, const int64_t *__restrict__ ctl) {
    if (ctl[CTL_DONE]) return;
    const W beta = scal[rr_new] / scal[rr_old];
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kCgThreads)
        p[i] = fma_w(beta, p[i], r[i]);
}

// p = r, partial = sum r.r
template <typename W>
__global__ __launch_bounds__(kCgThreads) void k_copy_dot(int64_t n,
                                                         const W *__restrict__ r,
                                                         W *__restrict__ p,
                                                         W *__restrict__ partial) {
    W sq = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kCgThreads) {
        const W v = r[i];
        p[i] = v;
        sq = fma_w(v, v, sq);
    }
    const W t = block_sum(sq);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Inner start: r32 = r64 / ||r64||, p32 = r32, d32 = 0, partial = sum r32^2
__global__ __launch_bounds__(kCgThreads) void k_start_inner(
    int64_t n, const double *__restrict__ scal64, int rr,
    const double *__restrict__ r64, float *__restrict__ r32,
    float *__restrict__ p32, float *__restrict__ d32, float *__restrict__ partial) {
    const double inv = 1.0 / sqrt(scal64[rr]);
    float sq = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kCgThreads) {
        const float v = static_cast<float>(r64[i] * inv);
        r32[i] = v;
        p32[i] = v;
        d32[i] = 0.0f;
        sq = __builtin_fmaf(v, v, sq);
    }
    const float t = block_sum(sq);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// x64 += ||r64|| * d32
__global__ __launch_bounds__(kCgThreads) void k_add_correction(
    int64_t n, const double *__restrict__ scal64, int rr,
    const float *__restrict__ d32, double *__restrict__ x64) {
    const double nrm = sqrt(scal64[rr]);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kCgThreads)
        x64[i] = __builtin_fma(nrm, static_cast<double>(d32[i]), x64[i]);
}

// x64 += double(xl), xl = 0  (refinement step of the reliable-update CG)
__global__ __launch_bounds__(kCgThreads) void k_fold_correction(
    int64_t n, float *__restrict__ xl, double *__restrict__ x64) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kCgThreads) {
        x64[i] += static_cast<double>(xl[i]);
        xl[i] = 0.0f;
    }
}

// r32 = float(r64), partial = sum r32^2
__global__ __launch_bounds__(kCgThreads) void k_replace_residual(
    int64_t n, const double *__restrict__ r64, float *__restrict__ r32,
    float *__restrict__ partial) {
    float sq = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kCgThreads) {
        const float v = static_cast<float>(r64[i]);
        r32[i] = v;
        sq = __builtin_fmaf(v, v, sq);
    }
    const float t = block_sum(sq);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// r32 = float(r64), p32 = r32, xl = 0, partial = sum r32^2
__global__ __launch_bounds__(kCgThreads) void k_start_reliable(
    int64_t n, const double *__restrict__ r64, float *__restrict__ r32,
    float *__restrict__ p32, float *__restrict__ xl, float *__restrict__ partial) {
    float sq = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kCgThreads) {
        const float v = static_cast<float>(r64[i]);
        r32[i] = v;
        p32[i] = v;
        xl[i] = 0.0f;
        sq = __builtin_fmaf(v, v, sq);
    }
    const float t = block_sum(sq);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void k_clear_done(int64_t *ctl) {
    if (threadIdx.x == 0) {
        ctl[CTL_DONE] = 0;
        ctl[CTL_BREAK] = 0;
    }
}

__global__ void k_to_slots(int64_t npad, int64_t n_rows,
                           const int32_t *__restrict__ slot_row,
                           const double *__restrict__ v, double *__restrict__ out) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    const int32_t r = slot_row[s];
    out[s] = r < n_rows ? v[r] : 0.0;
}

__global__ void k_from_slots(int64_t npad, int64_t n_rows,
                             const int32_t *__restrict__ slot_row,
                             const double *__restrict__ v, double *__restrict__ out) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    const int32_t r = slot_row[s];
    if (r < n_rows) out[r] = v[s];
}

__global__ void k_zero_ctl(int64_t *ctl) {
    if (threadIdx.x < 4) ctl[threadIdx.x] = 0;
}

// ------------------------------------------------------------ matrix copies
__global__ void k_col_slot(int64_t m, const int32_t *__restrict__ col,
                           const int32_t *__restrict__ perm,
                           int32_t *__restrict__ out) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < m) out[e] = perm[col[e]];
}

// rel[e] = col_slot[e] - s for the elements of slot s (0 for padding);
// flags[block] = 1 when one of the block's offsets does not fit int16.
__global__ __launch_bounds__(256) void k_col_rel(int64_t npad, int64_t C,
                                                 const int32_t *__restrict__ chunk_ptr,
                                                 const int32_t *__restrict__ chunk_len,
                                                 const int32_t *__restrict__ col_slot,
                                                 const int32_t *__restrict__ src,
                                                 int16_t *__restrict__ rel,
                                                 int32_t *__restrict__ flags) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int bad = 0;
    if (s < npad) {
        const int64_t c = s / C;
        const int64_t base = static_cast<int64_t>(chunk_ptr[c]) + (s - c * C);
        for (int j = 0; j < chunk_len[c]; ++j) {
            const int64_t o = base + static_cast<int64_t>(j) * C;
            const int64_t off = src[o] >= 0 ? col_slot[o] - s : 0;
            bad |= off < -32768 || off > 32767;
            rel[o] = static_cast<int16_t>(off);
        }
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) flags[blockIdx.x] = bad;
}

// out[0] = OR of flags[0..m) (one workgroup).
__global__ __launch_bounds__(1024) void k_or_flags(const int32_t *__restrict__ flags, int64_t m,
                                                   int64_t *__restrict__ out) {
    int any = 0;
    for (int64_t i = threadIdx.x; i < m; i += 1024) any |= flags[i];
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) out[0] = any;
}

// kTab8 tables (C = 64: one wavefront per chunk, lane = slot).  A chunk's
// table lists the distinct offsets of its elements in order of first use
// (element j before j + 1, lower lane first), padded with 0; the index bytes
// go to the word layout of apply_slot (bytes past the chunk's length are 0).
// bad[chunk] = 1 when the chunk uses more than kColTable offsets (the coding
// is then not used).
__global__ __launch_bounds__(256) void k_col_tab(int64_t n_chunks,
                                                 const int32_t *__restrict__ chunk_ptr,
                                                 const int32_t *__restrict__ chunk_len,
                                                 const int16_t *__restrict__ rel,
                                                 const int64_t *__restrict__ idx_row,
                                                 int16_t *__restrict__ tab,
                                                 uint8_t *__restrict__ idx_bytes,
                                                 int32_t *__restrict__ bad) {
    const int64_t c = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    if (c >= n_chunks) return;  // whole wavefronts
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int64_t base = static_cast<int64_t>(chunk_ptr[c]) + lane;
    const int len = chunk_len[c];
    uint8_t *mine = idx_bytes + (idx_row[c] * kWave + lane) * 4;  // byte j%4 of row j/4
    int entry = 0;  // lane t holds table entry t
    int count = 0, over = 0;
    for (int j = 0; j < len && !over; ++j) {
        const int off = rel[base + static_cast<int64_t>(j) * kWave];
        int pos = -1;
        for (int t = 0; t < count; ++t)
            if (__shfl(entry, t) == off) pos = t;
        for (;;) {
            const unsigned long long miss = __ballot(pos < 0);
            if (miss == 0) break;
            if (count == kColTable) {
                over = 1;
                break;
            }
            const int v = __shfl(off, __ffsll(static_cast<long long>(miss)) - 1);
            if (lane == count) entry = v;
            if (off == v) pos = count;
            ++count;
        }
        if (!over) mine[(j >> 2) * kWave * 4 + (j & 3)] = static_cast<uint8_t>(pos);
    }
    for (int j = len; j & 3; ++j) mine[(j >> 2) * kWave * 4 + (j & 3)] = 0;
    if (lane < kColTable) tab[c * kColTable + lane] = static_cast<int16_t>(entry);
    if (lane == 0) bad[c] = over;
}

__global__ void k_round_f32(int64_t m, const double *__restrict__ v,
                            float *__restrict__ out) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < m) out[e] = static_cast<float>(v[e]);
}

// binary64 -> binary32 -> bfloat16, both roundings to nearest even
__device__ __forceinline__ uint16_t to_bf16(double v) {
    const uint32_t u = __float_as_uint(static_cast<float>(v));
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__global__ void k_round_bf16(int64_t m, const double *__restrict__ v,
                             uint16_t *__restrict__ out) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < m) out[e] = to_bf16(v[e]);
}

// Per slot: the stored diagonal entry goes to diag[s] (binary32) and is
// zeroed in the bfloat16 copy; every other element is rounded to bfloat16.
__global__ void k_split_bf16(int64_t npad, int64_t C,
                             const int32_t *__restrict__ chunk_ptr,
                             const int32_t *__restrict__ chunk_len,
                             const int32_t *__restrict__ col_slot,
                             const double *__restrict__ val,
                             const int32_t *__restrict__ src,
                             uint16_t *__restrict__ out, float *__restrict__ diag) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    const int64_t c = s / C;
    const int64_t base = static_cast<int64_t>(chunk_ptr[c]) + (s - c * C);
    float d = 0.0f;
    for (int j = 0; j < chunk_len[c]; ++j) {
        const int64_t o = base + static_cast<int64_t>(j) * C;
        if (src[o] >= 0 && col_slot[o] == s) {
            d = static_cast<float>(val[o]);
            out[o] = 0;
        } else {
            out[o] = to_bf16(val[o]);
        }
    }
    diag[s] = d;
}

// ------------------------------------------------------------ host driver
struct Mat {
    int64_t n, C;
    const int32_t *chunk_ptr, *chunk_len, *col;
    const int16_t *rel;   // 16-bit column offsets from the slot, or nullptr
    const uint32_t *idx;     // kTab8 index words (with idx_row and tab), or nullptr
    const int64_t *idx_row;
    const int16_t *tab;
    int mode() const { return idx ? kTab8 : rel ? kRel16 : kCol32; }
    int32_t coding() const {
        return idx ? SMAX_HIP_COLUMNS_8 : rel ? SMAX_HIP_COLUMNS_16 : SMAX_HIP_COLUMNS_32;
    }
};

// q = A p with partial dots, through the matrix's column coding.
template <typename V, typename W, bool DIAG>
static void launch_spmv_dot(smax_hip_ctx ctx, dim3 sg, const Mat &M, const V *val,
                            const float *diag, const W *p, W *q, W *partial,
                            const int64_t *ctl) {
#define SMAX_CG_SPMV(MODE)                                                                   \
    hipLaunchKernelGGL((k_spmv_dot<V, W, DIAG, MODE>), sg, dim3(kCgThreads), 0, ctx->stream, \
                       M.n, M.C, M.chunk_ptr, M.chunk_len, M.col, M.rel, M.idx, M.idx_row,     \
                       M.tab, val, diag, p, q, partial, ctl)
    switch (M.mode()) {
    case kTab8: SMAX_CG_SPMV(kTab8); break;
    case kRel16: SMAX_CG_SPMV(kRel16); break;
    default: SMAX_CG_SPMV(kCol32); break;
    }
#undef SMAX_CG_SPMV
}

struct Work {
    smax_hip_ctx ctx;
    std::vector<void *> ptrs;
    explicit Work(smax_hip_ctx c) : ctx(c) {}
    ~Work() {
        (void)hipStreamSynchronize(ctx->stream);
        for (void *p : ptrs) dev_free(p);
    }
    template <typename T> int get(T **p, int64_t n) {
        void *q = nullptr;
        int rc = dev_alloc(ctx, &q, sizeof(T) * static_cast<size_t>(n), "cg");
        if (rc) return rc;
        ptrs.push_back(q);
        *p = static_cast<T *>(q);
        return SMAX_HIP_SUCCESS;
    }
};

template <typename T>
static int read_words(smax_hip_ctx ctx, const T *d, T *h, int n) {
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(ctx->host_word, d, sizeof(T) * n,
                                      hipMemcpyDeviceToHost, ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) h[i] = reinterpret_cast<T *>(ctx->host_word)[i];
    return SMAX_HIP_SUCCESS;
}

static dim3 upd_grid(int64_t n) {
    int64_t g = ceil_div(n, kCgThreads);
    return dim3(static_cast<unsigned>(g < kUpdBlocks ? (g > 0 ? g : 1) : kUpdBlocks));
}

// CG iterations from the current (x, r, p) with r.r in scal[0].  Stops once
// sqrt(r.r) <= stop or after max_it iterations; the host looks at the device
// every `batch` iterations.  Returns the iterations run and the final r.r.
template <typename V, typename W, bool DIAG>
static int cg_loop(smax_hip_ctx ctx, const Mat &A, const V *val, const float *diag,
                   W *x, W *r, W *p, W *q, W *partial, W *scal, int64_t *ctl,
                   double stop, int64_t max_it, int64_t batch, int64_t *iters,
                   double *rr_out) {
    const int64_t n = A.n;
    const dim3 sg = grid_for(n, kCgThreads), ug = upd_grid(n);
    const W stop2 = static_cast<W>(stop * stop);
    hipLaunchKernelGGL(k_zero_ctl, dim3(1), dim3(64), 0, ctx->stream, ctl);
    int64_t h[2] = {0, 0};
    W rr = 0;
    int rc = read_words(ctx, scal, &rr, 1);
    if (rc) return rc;
    int64_t done_it = 0;
    bool done = !(static_cast<double>(rr) > stop * stop);
    int par = 0;  // scal[par] holds the current r.r
    while (!done && done_it < max_it) {
        const int64_t nb = std::min<int64_t>(batch, max_it - done_it);
        for (int64_t b = 0, pp = par; b < nb; ++b, pp = 1 - pp) {
            const int cur = static_cast<int>(pp), nxt = 1 - cur;
            launch_spmv_dot<V, W, DIAG>(ctx, sg, A, val, diag, p, q, partial, ctl);
            hipLaunchKernelGGL((k_reduce_pq<W>), dim3(1), dim3(kReduceThreads), 0,
                               ctx->stream, partial, static_cast<int64_t>(sg.x),
                               scal + 2, ctl);
            hipLaunchKernelGGL((k_update_xr<W>), ug, dim3(kCgThreads), 0, ctx->stream,
                               n, scal, cur, 2, x, r, p, q, partial, ctl);
            hipLaunchKernelGGL((k_reduce<W>), dim3(1), dim3(kReduceThreads), 0,
                               ctx->stream, partial, static_cast<int64_t>(ug.x),
                               scal + nxt, ctl, stop2);
            hipLaunchKernelGGL((k_update_p<W>), ug, dim3(kCgThreads), 0, ctx->stream,
                               n, scal, nxt, cur, r, p, ctl);
        }
        SMAX_HIP_LAUNCHED(ctx, "cg iteration");
        rc = read_words(ctx, ctl, h, 2);
        if (rc) return rc;
        done_it = h[CTL_ITERS];
        par = static_cast<int>(done_it & 1);
        rc = read_words(ctx, scal + par, &rr, 1);
        if (rc) return rc;
        done = h[CTL_DONE] != 0 || !(rr == rr);
    }
    *iters = done_it;
    *rr_out = static_cast<double>(rr);
    // leave the current r.r in scal[0] for the caller
    if (par == 1)
        SMAX_HIP_CALL(ctx, hipMemcpyAsync(scal, scal + 1, sizeof(W),
                                          hipMemcpyDeviceToDevice, ctx->stream));
    return SMAX_HIP_SUCCESS;
}

// FP64 residual r = b - A x; ||r||^2 lands in scal[idx] and on the host.
static int residual(smax_hip_ctx ctx, const Mat &A, const double *val,
                    const double *b, const double *x, double *r, double *partial,
                    double *scal, int idx, double *rr) {
    const dim3 sg = grid_for(A.n, kCgThreads);
#define SMAX_CG_RESIDUAL(MODE)                                                             \
    hipLaunchKernelGGL(k_residual<MODE>, sg, dim3(kCgThreads), 0, ctx->stream, A.n, A.C,   \
                       A.chunk_ptr, A.chunk_len, A.col, A.rel, A.idx, A.idx_row, A.tab, val, \
                       b, x, r, partial)
    switch (A.mode()) {
    case kTab8: SMAX_CG_RESIDUAL(kTab8); break;
    case kRel16: SMAX_CG_RESIDUAL(kRel16); break;
    default: SMAX_CG_RESIDUAL(kCol32); break;
    }
#undef SMAX_CG_RESIDUAL
    hipLaunchKernelGGL((k_reduce<double>), dim3(1), dim3(kReduceThreads), 0, ctx->stream,
                       partial, static_cast<int64_t>(sg.x), scal + idx, nullptr, -1.0);
    SMAX_HIP_LAUNCHED(ctx, "k_residual");
    return read_words(ctx, scal + idx, rr, 1);
}

// ||v||^2 of a device vector into scal[idx] (and the host).
template <typename W>
static int norm2(smax_hip_ctx ctx, int64_t n, const W *v, W *tmp, W *partial, W *scal,
                 int idx, W *out) {
    hipLaunchKernelGGL((k_copy_dot<W>), upd_grid(n), dim3(kCgThreads), 0, ctx->stream,
                       n, v, tmp, partial);
    hipLaunchKernelGGL((k_reduce<W>), dim3(1), dim3(kReduceThreads), 0, ctx->stream,
                       partial, static_cast<int64_t>(upd_grid(n).x), scal + idx,
                       nullptr, W(-1));
    SMAX_HIP_LAUNCHED(ctx, "norm2");
    return read_words(ctx, scal + idx, out, 1);
}

static int prepare(smax_hip_ctx ctx, smax_hip_scs A, int lo_type, bool want_diag,
                   double *setup_s) {
    const auto t0 = std::chrono::steady_clock::now();
    const smax_hip_scs_info &in = A->info;
    const int64_t m = in.n_elements > 0 ? in.n_elements : 1;
    int rc = 0;
    if (in.sigma > 1 && !A->col_slot) {
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->col_slot), 4 * m, "cg col_slot");
        if (rc) return rc;
        hipLaunchKernelGGL(k_col_slot, grid_for(m, 256), dim3(256), 0, ctx->stream,
                           in.n_elements, in.col, in.perm, A->col_slot);
        SMAX_HIP_LAUNCHED(ctx, "k_col_slot");
    }
    const int32_t *cs = A->col_slot ? A->col_slot : in.col;
    if (A->col_rel_state == 0) {
        // 16-bit slot-relative columns halve the index stream of every SpMV
        // when the matrix's bandwidth in slot space allows (stencils do)
        const int64_t nb = grid_for(in.n_rows_padded, 256).x;
        int32_t *flags = nullptr;
        int64_t *word = nullptr;
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->col_rel), 2 * m, "cg col_rel");
        if (!rc) rc = dev_alloc(ctx, reinterpret_cast<void **>(&flags), 4 * nb, "cg col_rel");
        if (!rc) rc = dev_alloc(ctx, reinterpret_cast<void **>(&word), 8, "cg col_rel");
        int64_t any = 1;
        if (!rc) {
            hipLaunchKernelGGL(k_col_rel, dim3(static_cast<unsigned>(nb)), dim3(256), 0,
                               ctx->stream, in.n_rows_padded, in.C, in.chunk_ptr,
                               in.chunk_lengths, cs, in.src_index, A->col_rel, flags);
            hipLaunchKernelGGL(k_or_flags, dim3(1), dim3(1024), 0, ctx->stream, flags, nb,
                               word);
            rc = check_launch(ctx, "k_col_rel");
            if (!rc) rc = read_words(ctx, word, &any, 1);
        }
        dev_free(flags);
        dev_free(word);
        if (rc) {
            dev_free(A->col_rel);
            A->col_rel = nullptr;
            return rc;
        }
        if (any) {
            dev_free(A->col_rel);
            A->col_rel = nullptr;
        }
        A->col_rel_state = any ? -1 : 1;
    }
    if (A->col_tab_state == 0) {
        // one-byte indices into per-chunk offset tables: a quarter of the
        // 32-bit index stream, when every chunk uses few distinct offsets
        // (a stencil's chunks use one per stencil point)
        A->col_tab_state = -1;
        if (A->col_rel_state == 1 && in.C == kWave && in.n_chunks > 0) {
            const int64_t nc = in.n_chunks;
            // word rows per chunk: ceil(len / 4); prefix sums on the host
            std::vector<int32_t> lens(static_cast<size_t>(nc));
            std::vector<int64_t> rows(static_cast<size_t>(nc));
            rc = smax_hip_memcpy_d2h(ctx, lens.data(), in.chunk_lengths, 4 * nc);
            if (rc) return rc;
            int64_t total = 0;
            for (int64_t k = 0; k < nc; ++k) {
                rows[k] = total;
                total += (lens[k] + 3) / 4;
            }
            int32_t *flags = nullptr;
            int64_t *word = nullptr;
            rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->col_idx),
                           4 * kWave * (total > 0 ? total : 1), "cg col_idx");
            if (!rc)
                rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->col_idx_row), 8 * nc,
                               "cg col_idx");
            if (!rc)
                rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->col_tab),
                               2 * kColTable * nc, "cg col_tab");
            if (!rc) rc = dev_alloc(ctx, reinterpret_cast<void **>(&flags), 4 * nc, "cg col_tab");
            if (!rc) rc = dev_alloc(ctx, reinterpret_cast<void **>(&word), 8, "cg col_tab");
            if (!rc)
                rc = smax_hip_memcpy_h2d(ctx, A->col_idx_row, rows.data(), 8 * nc);
            int64_t any = 1;
            if (!rc) {
                hipLaunchKernelGGL(k_col_tab, grid_for(nc * kWave, 256), dim3(256), 0,
                                   ctx->stream, nc, in.chunk_ptr, in.chunk_lengths,
                                   A->col_rel, A->col_idx_row, A->col_tab,
                                   reinterpret_cast<uint8_t *>(A->col_idx), flags);
                hipLaunchKernelGGL(k_or_flags, dim3(1), dim3(1024), 0, ctx->stream, flags,
                                   nc, word);
                rc = check_launch(ctx, "k_col_tab");
                if (!rc) rc = read_words(ctx, word, &any, 1);
            }
            dev_free(flags);
            dev_free(word);
            if (rc || any) {
                dev_free(A->col_idx);
                dev_free(A->col_idx_row);
                dev_free(A->col_tab);
                A->col_idx = nullptr;
                A->col_idx_row = nullptr;
                A->col_tab = nullptr;
                if (rc) {
                    A->col_tab_state = 0;
                    return rc;
                }
            } else {
                A->col_tab_state = 1;
                A->col_idx_rows = total;
            }
        }
    }
    const double *v64 = static_cast<const double *>(in.val);
    if (lo_type == SMAX_HIP_F32 && !A->val_f32) {
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->val_f32), 4 * m, "cg f32");
        if (rc) return rc;
        hipLaunchKernelGGL(k_round_f32, grid_for(m, 256), dim3(256), 0, ctx->stream,
                           in.n_elements, v64, A->val_f32);
        SMAX_HIP_LAUNCHED(ctx, "k_round_f32");
    }
    if (lo_type == SMAX_HIP_BF16 && (!A->val_bf16 || A->bf16_split != (want_diag ? 1 : 0))) {
        if (!A->val_bf16) {
            rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->val_bf16), 2 * m, "cg bf16");
            if (rc) return rc;
        }
        if (want_diag) {
            if (!A->diag_f32) {
                rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->diag_f32),
                               4 * in.n_rows_padded, "cg diag");
                if (rc) return rc;
            }
            hipLaunchKernelGGL(k_split_bf16, grid_for(in.n_rows_padded, 256), dim3(256),
                               0, ctx->stream, in.n_rows_padded, in.C, in.chunk_ptr,
                               in.chunk_lengths, cs, v64, in.src_index, A->val_bf16,
                               A->diag_f32);
            SMAX_HIP_LAUNCHED(ctx, "k_split_bf16");
        } else {
            hipLaunchKernelGGL(k_round_bf16, grid_for(m, 256), dim3(256), 0, ctx->stream,
                               in.n_elements, v64, A->val_bf16);
            SMAX_HIP_LAUNCHED(ctx, "k_round_bf16");
        }
        A->bf16_split = want_diag ? 1 : 0;
    }
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    *setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                   .count();
    return SMAX_HIP_SUCCESS;
}

// Mixed-precision CG with reliable updates: low-precision (V values, FP32
// vectors) CG iterations run in device-side batches until the recursive
// residual has dropped by `delta` since the last refinement step (or below
// the target); then the correction is folded into the FP64 solution, the
// FP64 residual b - A x replaces the low-precision one, and the iteration
// continues with its search direction p (beta from the replaced residual).
//
// Safeguards, for low-precision operators far from A: a cycle also ends
// after max_inner iterations or at a p.q <= 0 breakdown, and then restarts
// from the FP64 residual (p = r).  A refinement step that does not lower the
// FP64 residual is undone (x restored from `xbak`) and restarts the
// iteration; after two such steps in a row `*gave_up` hands the rest of the
// solve to restarted (defect-correction) refinement.
template <typename V, bool DIAG>
static int reliable_cg(smax_hip_ctx ctx, const Mat &M, const double *val64,
                       const V *val, const float *diag, const double *bs,
                       double *xs, double *xbak, double *r64, double *partial64,
                       double *scal64, float *xl, float *r, float *p, float *q,
                       float *partial32, float *scal32, int64_t *ctl, double bnorm,
                       const smax_hip_cg_params *prm, int64_t batch, int64_t max_inner,
                       int64_t *iters, int64_t *steps, double *rr_out,
                       bool *gave_up) {
    const int64_t n = M.n;
    const dim3 sg = grid_for(n, kCgThreads), ug = upd_grid(n);
    const double target = prm->rel_tol * bnorm;
    const double delta = prm->inner_rel_tol;
    double rr64 = 0;
    int rc = residual(ctx, M, val64, bs, xs, r64, partial64, scal64, 0, &rr64);
    if (rc) return rc;
    *rr_out = rr64;
    if (bnorm == 0 || std::sqrt(rr64) <= target) return SMAX_HIP_SUCCESS;
    hipLaunchKernelGGL(k_start_reliable, ug, dim3(kCgThreads), 0, ctx->stream, n, r64,
                       r, p, xl, partial32);
    hipLaunchKernelGGL((k_reduce<float>), dim3(1), dim3(kReduceThreads), 0, ctx->stream,
                       partial32, static_cast<int64_t>(ug.x), scal32 + 0, nullptr, -1.0f);
    hipLaunchKernelGGL(k_zero_ctl, dim3(1), dim3(64), 0, ctx->stream, ctl);
    SMAX_HIP_LAUNCHED(ctx, "reliable start");
    double max_r = std::sqrt(rr64);
    int64_t done_it = 0, fails = 0;
    while (done_it < prm->max_iters) {
        const double trig = std::max(delta * max_r, 0.5 * target);
        const float trig2 = static_cast<float>(trig * trig);
        const int64_t cycle_end = std::min(prm->max_iters, done_it + max_inner);
        int64_t h[3] = {0, 0, 0};
        bool stopped = false;
        while (!stopped && done_it < cycle_end) {
            const int64_t nb = std::min<int64_t>(batch, cycle_end - done_it);
            for (int64_t b = 0, pp = done_it & 1; b < nb; ++b, pp = 1 - pp) {
                const int cur = static_cast<int>(pp), nxt = 1 - cur;
                launch_spmv_dot<V, float, DIAG>(ctx, sg, M, val, diag, p, q, partial32,
                                                ctl);
                hipLaunchKernelGGL((k_reduce_pq<float>), dim3(1), dim3(kReduceThreads), 0,
                                   ctx->stream, partial32, static_cast<int64_t>(sg.x),
                                   scal32 + 2, ctl);
                hipLaunchKernelGGL((k_update_xr<float>), ug, dim3(kCgThreads), 0,
                                   ctx->stream, n, scal32, cur, 2, xl, r, p, q,
                                   partial32, ctl);
                hipLaunchKernelGGL((k_reduce<float>), dim3(1), dim3(kReduceThreads), 0,
                                   ctx->stream, partial32, static_cast<int64_t>(ug.x),
                                   scal32 + nxt, ctl, trig2);
                hipLaunchKernelGGL((k_update_p<float>), ug, dim3(kCgThreads), 0,
                                   ctx->stream, n, scal32, nxt, cur, r, p, ctl);
            }
            SMAX_HIP_LAUNCHED(ctx, "reliable iteration");
            rc = read_words(ctx, ctl, h, 3);
            if (rc) return rc;
            done_it = h[CTL_ITERS];
            stopped = h[CTL_DONE] != 0;
        }
        // Continue p only after a regular stop: the r.r trigger skipped the
        // p update of its iteration, which the refinement step completes.
        bool keep_p = stopped && h[CTL_BREAK] == 0;
        // refinement step in FP64
        ++*steps;
        const int par = static_cast<int>(done_it & 1);  // scal32[par]: newest r.r
        SMAX_HIP_CALL(ctx, hipMemcpyAsync(xbak, xs, sizeof(double) * n,
                                          hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(k_fold_correction, ug, dim3(kCgThreads), 0, ctx->stream, n,
                           xl, xs);
        rc = residual(ctx, M, val64, bs, xs, r64, partial64, scal64, 0, &rr64);
        if (rc) return rc;
        double nr = std::sqrt(rr64);
        if (nr <= target) {
            *rr_out = rr64;
            break;
        }
        if (nr < max_r) {
            fails = 0;
        } else {
            // undo the step: x and r back to the last refinement point
            SMAX_HIP_CALL(ctx, hipMemcpyAsync(xs, xbak, sizeof(double) * n,
                                              hipMemcpyDeviceToDevice, ctx->stream));
            rc = residual(ctx, M, val64, bs, xs, r64, partial64, scal64, 0, &rr64);
            if (rc) return rc;
            nr = std::sqrt(rr64);
            keep_p = false;
            if (++fails >= 2) {
                *rr_out = rr64;
                *gave_up = true;
                break;
            }
        }
        *rr_out = rr64;
        max_r = nr;
        if (keep_p)
            hipLaunchKernelGGL(k_replace_residual, ug, dim3(kCgThreads), 0, ctx->stream,
                               n, r64, r, partial32);
        else
            hipLaunchKernelGGL(k_start_reliable, ug, dim3(kCgThreads), 0, ctx->stream,
                               n, r64, r, p, xl, partial32);
        hipLaunchKernelGGL((k_reduce<float>), dim3(1), dim3(kReduceThreads), 0, ctx->stream,
                           partial32, static_cast<int64_t>(ug.x), scal32 + par, nullptr,
                           -1.0f);
        hipLaunchKernelGGL(k_clear_done, dim3(1), dim3(64), 0, ctx->stream, ctl);
        if (keep_p)
            hipLaunchKernelGGL((k_update_p<float>), ug, dim3(kCgThreads), 0, ctx->stream,
                               n, scal32, par, 1 - par, r, p, ctl);
        SMAX_HIP_LAUNCHED(ctx, "refinement step");
    }
    *iters = done_it;
    return SMAX_HIP_SUCCESS;
}

static int check_solver_args(smax_hip_ctx ctx, const char *fn, smax_hip_scs A,
                             const double *b, double *x,
                             const smax_hip_cg_params *prm,
                             smax_hip_cg_report *rep) {
    SMAX_HIP_REQUIRE(ctx, A && b && x && prm && rep, "%s: null argument", fn);
    SMAX_HIP_REQUIRE(ctx, !A->from_sell,
                     "%s: the matrix was built from SELL arrays; build it with "
                     "smax_hip_scs_from_crs", fn);
    SMAX_HIP_REQUIRE(ctx, A->info.val_bytes == 8,
                     "%s: the matrix must hold FP64 values", fn);
    SMAX_HIP_REQUIRE(ctx, A->info.n_rows == A->info.n_cols,
                     "%s: the matrix must be square", fn);
    SMAX_HIP_REQUIRE(ctx, prm->rel_tol > 0 && prm->max_iters > 0,
                     "%s: rel_tol and max_iters must be positive", fn);
    return SMAX_HIP_SUCCESS;
}

// The solver matrix in slot space with the narrowest column coding built for
// it, limited by params->column_coding (SMAX_HIP_COLUMNS_*).
static Mat slot_matrix(smax_hip_scs A, int32_t coding) {
    const smax_hip_scs_info &in = A->info;
    Mat M{in.n_rows_padded, in.C, in.chunk_ptr, in.chunk_lengths,
          A->col_slot ? A->col_slot : in.col, nullptr, nullptr, nullptr, nullptr};
    if (coding != SMAX_HIP_COLUMNS_32) M.rel = A->col_rel;
    if ((coding == SMAX_HIP_COLUMNS_AUTO || coding == SMAX_HIP_COLUMNS_8) && A->col_idx &&
        in.n_rows_padded % kWave == 0) {
        M.idx = A->col_idx;
        M.idx_row = A->col_idx_row;
        M.tab = A->col_tab;
    }
    return M;
}

}  // namespace smaxhip

using namespace smaxhip;

extern "C" {

int smax_hip_cg_f64(smax_hip_ctx ctx, smax_hip_scs A, const double *b,
                    double *x, const smax_hip_cg_params *prm,
                    smax_hip_cg_report *rep) {
    SMAX_HIP_CHECK_CTX(ctx);
    int rc = check_solver_args(ctx, "smax_hip_cg_f64", A, b, x, prm, rep);
    if (rc) return rc;
    *rep = smax_hip_cg_report{};
    rc = prepare(ctx, A, -1, false, &rep->setup_seconds);
    if (rc) return rc;
    const smax_hip_scs_info &in = A->info;
    const Mat M = slot_matrix(A, prm->column_coding);
    rep->column_coding = M.coding();
    const double *val = static_cast<const double *>(in.val);
    const int64_t n = M.n, batch = prm->check_every > 0 ? prm->check_every : 25;
    Work w(ctx);
    double *bs, *xs, *r, *p, *q, *partial, *scal;
    int64_t *ctl;
    rc = w.get(&bs, n) | w.get(&xs, n) | w.get(&r, n) | w.get(&p, n) |
         w.get(&q, n) | w.get(&partial, ceil_div(n, kCgThreads) + kUpdBlocks) |
         w.get(&scal, 8) | w.get(&ctl, 4);
    if (rc) return SMAX_HIP_ERROR_OUT_OF_MEMORY;
    const auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_to_slots, grid_for(n, 256), dim3(256), 0, ctx->stream, n,
                       in.n_rows, in.slot_row, b, bs);
    hipLaunchKernelGGL(k_to_slots, grid_for(n, 256), dim3(256), 0, ctx->stream, n,
                       in.n_rows, in.slot_row, x, xs);
    double bb = 0, rr = 0;
    rc = norm2(ctx, n, bs, q, partial, scal, 3, &bb);
    if (rc) return rc;
    const double bnorm = std::sqrt(bb), stop = prm->rel_tol * bnorm;
    int64_t total = 0;
    for (int restart = 0; restart < 4; ++restart) {
        // (re)start from the true residual
        rc = residual(ctx, M, val, bs, xs, r, partial, scal, 0, &rr);
        if (rc) return rc;
        if (bnorm == 0 || std::sqrt(rr) <= stop || total >= prm->max_iters) break;
        hipLaunchKernelGGL((k_copy_dot<double>), upd_grid(n), dim3(kCgThreads), 0,
                           ctx->stream, n, r, p, partial);
        int64_t it = 0;
        double rr_rec = 0;
        rc = cg_loop<double, double, false>(ctx, M, val, nullptr, xs, r, p, q, partial,
                                            scal, ctl, stop, prm->max_iters - total,
                                            batch, &it, &rr_rec);
        if (rc) return rc;
        total += it;
    }
    hipLaunchKernelGGL(k_from_slots, grid_for(n, 256), dim3(256), 0, ctx->stream, n,
                       in.n_rows, in.slot_row, xs, x);
    SMAX_HIP_LAUNCHED(ctx, "k_from_slots");
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    rep->seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rep->iterations = total;
    rep->rel_residual = bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
    if (!(rep->rel_residual <= prm->rel_tol))
        return fail(ctx, SMAX_HIP_ERROR_NOT_CONVERGED,
                    "smax_hip_cg_f64: relative residual %.3e after %lld iterations",
                    rep->rel_residual, (long long)total);
    return SMAX_HIP_SUCCESS;
}

int smax_hip_ircg(smax_hip_ctx ctx, smax_hip_scs A, const double *b, double *x,
                  const smax_hip_cg_params *prm, smax_hip_cg_report *rep) {
    SMAX_HIP_CHECK_CTX(ctx);
    int rc = check_solver_args(ctx, "smax_hip_ircg", A, b, x, prm, rep);
    if (rc) return rc;
    const int lo = prm->inner_value_type;
    SMAX_HIP_REQUIRE(ctx, lo == SMAX_HIP_F32 || lo == SMAX_HIP_BF16,
                     "smax_hip_ircg: inner_value_type must be F32 or BF16");
    SMAX_HIP_REQUIRE(ctx, prm->inner_rel_tol > 0 && prm->inner_rel_tol < 1,
                     "smax_hip_ircg: inner_rel_tol must lie in (0, 1)");
    *rep = smax_hip_cg_report{};
    // the FP32 diagonal side array applies to BF16 storage only
    const bool diag = prm->keep_diag_f32 != 0 && lo == SMAX_HIP_BF16;
    rc = prepare(ctx, A, lo, diag, &rep->setup_seconds);
    if (rc) return rc;
    const smax_hip_scs_info &in = A->info;
    const Mat M = slot_matrix(A, prm->column_coding);
    rep->column_coding = M.coding();

    const double *val = static_cast<const double *>(in.val);
    const int64_t n = M.n, batch = prm->check_every > 0 ? prm->check_every : 25;
    const int64_t max_inner = prm->max_inner > 0 ? prm->max_inner : prm->max_iters;
    Work w(ctx);
    double *bs, *xs, *xbak, *r64, *partial64, *scal64;
    float *d, *r, *p, *q, *partial32, *scal32;
    int64_t *ctl;
    const int64_t np = ceil_div(n, kCgThreads) + kUpdBlocks;
    rc = w.get(&bs, n) | w.get(&xs, n) | w.get(&xbak, n) | w.get(&r64, n) |
         w.get(&partial64, np) |
         w.get(&scal64, 8) | w.get(&d, n) | w.get(&r, n) | w.get(&p, n) |
         w.get(&q, n) | w.get(&partial32, np) | w.get(&scal32, 8) | w.get(&ctl, 4);
    if (rc) return SMAX_HIP_ERROR_OUT_OF_MEMORY;
    const auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_to_slots, grid_for(n, 256), dim3(256), 0, ctx->stream, n,
                       in.n_rows, in.slot_row, b, bs);
    hipLaunchKernelGGL(k_to_slots, grid_for(n, 256), dim3(256), 0, ctx->stream, n,
                       in.n_rows, in.slot_row, x, xs);
    double bb = 0, rr = 0;
    rc = norm2(ctx, n, bs, r64, partial64, scal64, 3, &bb);
    if (rc) return rc;
    const double bnorm = std::sqrt(bb);
    int64_t total = 0, outer = 0;
    bool fallback = false;
    if (!prm->restart_inner) {
        if (lo == SMAX_HIP_F32)
            rc = reliable_cg<float, false>(ctx, M, val, A->val_f32, nullptr, bs, xs, xbak,
                                           r64, partial64, scal64, d, r, p, q, partial32,
                                           scal32, ctl, bnorm, prm, batch, max_inner,
                                           &total, &outer, &rr, &fallback);
        else if (diag)
            rc = reliable_cg<uint16_t, true>(ctx, M, val, A->val_bf16, A->diag_f32, bs,
                                             xs, xbak, r64, partial64, scal64, d, r, p,
                                             q, partial32, scal32, ctl, bnorm, prm,
                                             batch, max_inner, &total, &outer, &rr,
                                             &fallback);
        else
            rc = reliable_cg<uint16_t, false>(ctx, M, val, A->val_bf16, nullptr, bs, xs,
                                              xbak, r64, partial64, scal64, d, r, p, q,
                                              partial32, scal32, ctl, bnorm, prm, batch,
                                              max_inner, &total, &outer, &rr, &fallback);
        if (rc) return rc;
    } else {
        rc = residual(ctx, M, val, bs, xs, r64, partial64, scal64, 0, &rr);
        if (rc) return rc;
    }
    // Restarted refinement: each step solves A d = r (r / ||r|| in FP32) with
    // low-precision CG, x += ||r|| d in FP64; a step that does not lower the
    // FP64 residual is undone and ends the solve.  r64 / scal64[0] / rr hold
    // the residual of the current x on entry to every step.
    while (prm->restart_inner || fallback) {
        const double rel = bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
        if (rel <= prm->rel_tol || total >= prm->max_iters) break;
        ++outer;
        hipLaunchKernelGGL(k_start_inner, upd_grid(n), dim3(kCgThreads), 0,
                           ctx->stream, n, scal64, 0, r64, r, p, d, partial32);
        hipLaunchKernelGGL((k_reduce<float>), dim3(1), dim3(kReduceThreads), 0, ctx->stream,
                           partial32, static_cast<int64_t>(upd_grid(n).x), scal32 + 0,
                           nullptr, -1.0f);
        SMAX_HIP_LAUNCHED(ctx, "k_start_inner");
        // Solve no deeper than the outer target needs (with a 1/4 margin).
        const double inner_tol =
            std::max(prm->inner_rel_tol, 0.25 * prm->rel_tol / rel);
        int64_t it = 0;
        double rr_in = 0;
        const int64_t cap = std::min<int64_t>(max_inner, prm->max_iters - total);
        if (lo == SMAX_HIP_F32)
            rc = cg_loop<float, float, false>(ctx, M, A->val_f32, nullptr, d, r, p, q,
                                              partial32, scal32, ctl, inner_tol, cap,
                                              batch, &it, &rr_in);
        else if (diag)
            rc = cg_loop<uint16_t, float, true>(ctx, M, A->val_bf16, A->diag_f32, d, r,
                                                p, q, partial32, scal32, ctl, inner_tol,
                                                cap, batch, &it, &rr_in);
        else
            rc = cg_loop<uint16_t, float, false>(ctx, M, A->val_bf16, nullptr, d, r, p,
                                                 q, partial32, scal32, ctl, inner_tol,
                                                 cap, batch, &it, &rr_in);
        if (rc) return rc;
        total += it;
        SMAX_HIP_CALL(ctx, hipMemcpyAsync(xbak, xs, sizeof(double) * n,
                                          hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(k_add_correction, upd_grid(n), dim3(kCgThreads), 0,
                           ctx->stream, n, scal64, 0, d, xs);
        SMAX_HIP_LAUNCHED(ctx, "k_add_correction");
        double rr_new = 0;
        rc = residual(ctx, M, val, bs, xs, r64, partial64, scal64, 0, &rr_new);
        if (rc) return rc;
        if (!(rr_new < rr)) {
            SMAX_HIP_CALL(ctx, hipMemcpyAsync(xs, xbak, sizeof(double) * n,
                                              hipMemcpyDeviceToDevice, ctx->stream));
            break;
        }
        rr = rr_new;
    }
    hipLaunchKernelGGL(k_from_slots, grid_for(n, 256), dim3(256), 0, ctx->stream, n,
                       in.n_rows, in.slot_row, xs, x);
    SMAX_HIP_LAUNCHED(ctx, "k_from_slots");
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    rep->seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rep->iterations = total;
    rep->outer_iterations = outer;
    rep->rel_residual = bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
    if (!(rep->rel_residual <= prm->rel_tol))
        return fail(ctx, SMAX_HIP_ERROR_NOT_CONVERGED,
                    "smax_hip_ircg: relative residual %.3e after %lld outer / %lld "
                    "inner iterations",
                    rep->rel_residual, (long long)outer, (long long)total);
    return SMAX_HIP_SUCCESS;
}

}  // extern "C"
