// SELL-C-sigma SpMV and SpMM for gfx950.
//
// Layout (src/utils/convert_formats.hpp:16-132): chunk c holds C slots; its
// element j of slot i sits at chunk_ptr[c] + j*C + i, so for C = 64 one
// wavefront reads a whole 64-slot column of a chunk with one coalesced
// load per array.  Every slot is a sequential fma chain over j, exactly the
// loop of the reference (spmv_cpu_scs_impl.hpp:27-41), so results are
// bit-identical to it; padding elements (val 0, col 0) add an exact zero.
//
// The chains are unrolled: U column indices and values are loaded
// (non-temporally, they are streamed once) before the U gathers of x and the
// U dependent fmas, so each lane keeps 2U loads in flight.
#include <algorithm>
#include <cmath>
#include <functional>
#include <utility>
#include <vector>

#include "common.hpp"
#include "convert_internal.hpp"

namespace smaxhip {

constexpr int kScsThreads = 256;
constexpr int kScsUnroll = 8;       // generic-C chains
// C = 64 chains, elements per group (measured on the 4M-row band matrix):
// FP64 24 (0.697 / 0.641 ms at sigma = 4096 / 65536; 16: 0.707 / 0.659, with
// 128 VGPRs against 90 -- four waves per SIMD instead of five, each with more
// gathers in flight; 28 and 32 leave two waves and lose at sigma = 4096),
// FP32 16 (24: 0.519 vs 0.493 ms), BF16 values 8 (16: 0.52 vs 0.48 ms).
template <int VT>
constexpr int kScsPipeUnroll = VT == SMAX_HIP_BF16 ? 8 : VT == SMAX_HIP_F32 ? 16 : 24;

// Layouts whose chunks are at most kShortChain long (stencils: the 512^2
// Poisson matrix has 5) run each chain as one predicated group: all of a
// lane's column/value loads, then all its gathers, then the fmas, two memory
// round trips in all.  The 512^2 Poisson SpMV after the caches are flushed:
// 7.16 -> 5.98 us kernel time; back to back 4.56 -> 4.68 us.
constexpr int kShortChain = 8;

// One slot's chain in groups of U: the next group's column indices and
// values are requested before this group's x gathers (software pipelining),
// so a long chain waits about one memory round trip per group instead of
// two.  The remainder runs element by element: a predicated group for it
// costs registers and measured slower on the 4M-row band matrix.
template <int VT, int U>
__device__ __forceinline__ typename vt<VT>::vec_t
slot_chain_pipe(const int32_t *__restrict__ col,
                const typename vt<VT>::val_t *__restrict__ val,
                const typename vt<VT>::vec_t *__restrict__ x, int64_t base,
                int64_t stride, int len) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    using V = typename T::val_t;
    A acc = 0;
    int j = 0;
    if (len >= U) {
        int32_t c0[U];
        V v0[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c0[u] = ld_stream(col + base + u * stride);
            v0[u] = ld_stream(val + base + u * stride);
        }
        for (; j + 2 * U <= len; j += U) {
            int32_t c1[U];
            V v1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c1[u] = ld_stream(col + base + (j + U + u) * stride);
                v1[u] = ld_stream(val + base + (j + U + u) * stride);
            }
            A xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[c0[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) acc = T::fma_(T::widen(v0[u]), xv[u], acc);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c0[u] = c1[u];
                v0[u] = v1[u];
            }
        }
        A xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[c0[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) acc = T::fma_(T::widen(v0[u]), xv[u], acc);
        j += U;
    }
    for (; j < len; ++j) {
        const int32_t c = ld_stream(col + base + j * stride);
        const auto v = ld_stream(val + base + j * stride);
        acc = T::fma_(T::widen(v), x[c], acc);
    }
    return acc;
}

// C == 64: one wavefront per chunk, one lane per slot.
//   REMAP: XCD-contiguous workgroup order (see xcd_block).
//   order: optional execution order, wave position -> chunk (-1: idle).
//   LEN:   each lane stops at its row's last stored element (slot_len); a
//          lane with padding then applies the reference's padding term
//          fma(0, x[0], acc) once, which equals applying it any number of
//          times (f(a) = fma(0, x0, a) satisfies f(f(a)) = f(a) for every a,
//          signed zeros and non-finite x0 included), so results stay
//          bit-identical to the full-chunk chain.
//   SHORT: every chunk is at most kShortChain long (single-group chains;
//          no tail waves).
//   tails (LEN only; t_slot null: none): the first t_blocks workgroups run
//          the handle's tail waves (smax_hip_scs_tails), the rest the chunks,
//          whose lanes longer than t_head[chunk] then leave at once.  Both
//          kinds of wave run the same chain code on their own arrays;
//          t_blocks is a multiple of 8, so the chunks' workgroups keep their
//          XCD alignment.
template <int VT, bool UNPERM, bool REMAP, bool LEN, bool SHORT>
__global__ __launch_bounds__(kScsThreads) void scs_spmv_w64(
    int64_t n_chunks, int64_t n_rows, const int32_t *__restrict__ chunk_ptr,
    const int32_t *__restrict__ chunk_len, const int32_t *__restrict__ slot_len,
    const int32_t *__restrict__ order, const int32_t *__restrict__ col,
    const typename vt<VT>::val_t *__restrict__ val,
    const int32_t *__restrict__ slot_row,
    const typename vt<VT>::vec_t *__restrict__ x,
    typename vt<VT>::vec_t *__restrict__ y, const int32_t *__restrict__ t_head,
    const int32_t *__restrict__ t_slot, const int64_t *__restrict__ t_ptr,
    const int32_t *__restrict__ t_col, const typename vt<VT>::val_t *__restrict__ t_val,
    int64_t t_blocks) {
    using T = vt<VT>;
    const int lane = threadIdx.x & (kWave - 1);
    const bool tails = !SHORT && LEN && t_slot != nullptr;
    int64_t slot, base;
    int clen, len;
    const int32_t *cp;
    const typename T::val_t *vp;
    if (tails && static_cast<int64_t>(blockIdx.x) < t_blocks) {
        const int64_t w = static_cast<int64_t>(blockIdx.x) * (kScsThreads / kWave) +
                          threadIdx.x / kWave;
        const int32_t ts = t_slot[w * kWave + lane];
        if (ts < 0) return;
        slot = ts;
        base = t_ptr[w] + lane;
        clen = chunk_len[ts / kWave];
        len = slot_len[ts];
        cp = t_col;
        vp = t_val;
    } else {
        const int64_t hb = tails ? blockIdx.x - t_blocks : blockIdx.x;
        const int64_t hn = tails ? gridDim.x - t_blocks : gridDim.x;
        const int64_t blk = REMAP ? xcd_block(hb, hn) : hb;
        const int64_t pos = blk * (kScsThreads / kWave) + (threadIdx.x / kWave);
        const int64_t chunk = order ? static_cast<int64_t>(order[pos]) : pos;
        if (chunk < 0 || chunk >= n_chunks) return;
        slot = chunk * kWave + lane;
        base = static_cast<int64_t>(chunk_ptr[chunk]) + lane;
        clen = chunk_len[chunk];
        len = LEN ? slot_len[slot] : clen;
        if (tails && len > t_head[chunk]) return;
        cp = col;
        vp = val;
    }
    typename T::vec_t acc = 0;
    if constexpr (SHORT) {
        int32_t c[kShortChain];
        typename T::val_t v[kShortChain];
#pragma unroll
        for (int u = 0; u < kShortChain; ++u)
            if (u < len) {
                c[u] = ld_stream(cp + base + u * kWave);
                v[u] = ld_stream(vp + base + u * kWave);
            }
        typename T::vec_t xv[kShortChain];
#pragma unroll
        for (int u = 0; u < kShortChain; ++u)
            if (u < len) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < kShortChain; ++u)
            if (u < len) acc = T::fma_(T::widen(v[u]), xv[u], acc);
    } else {
        acc = slot_chain_pipe<VT, kScsPipeUnroll<VT>>(cp, vp, x, base, kWave, len);
    }
    if (LEN && len < clen) acc = T::fma_(T::widen(typename T::val_t(0)), x[0], acc);
    if (UNPERM) {
        const int32_t r = slot_row[slot];
        if (r < n_rows) y[r] = acc;
    } else {
        y[slot] = acc;
    }
}

// Any C: one lane per slot, 256 / C chunks per workgroup (C <= 256) or one
// chunk over several workgroups; chunks in the optional execution `order`;
// LEN and the tail waves (C <= 256 only; their arrays are a SELL-64 whatever
// C is) as in scs_spmv_w64.
//   CS > 0: C fixed at compile time (layouts without tail waves): the chain
//   then addresses its elements with constant offsets and runs the C = 64
//   kernel's group size (4M-row band matrix, C = 32, FP64: sigma = 65536
//   0.692 -> 0.660 ms, 0.640 with the occupancy cap; sigma = N 1.093 ->
//   1.061 ms; BF16 values, sigma = 65536: 0.495 -> 0.477 ms).
template <int VT, bool UNPERM, bool LEN, bool REMAP, int CS>
__global__ __launch_bounds__(kScsThreads) void scs_spmv_generic(
    int64_t C, int64_t n_chunks, int64_t n_rows, int64_t cpb, int64_t bpc,
    const int32_t *__restrict__ chunk_ptr, const int32_t *__restrict__ chunk_len,
    const int32_t *__restrict__ slot_len, const int32_t *__restrict__ order,
    const int32_t *__restrict__ col,
    const typename vt<VT>::val_t *__restrict__ val,
    const int32_t *__restrict__ slot_row,
    const typename vt<VT>::vec_t *__restrict__ x,
    typename vt<VT>::vec_t *__restrict__ y, const int32_t *__restrict__ t_head,
    const int32_t *__restrict__ t_slot, const int64_t *__restrict__ t_ptr,
    const int32_t *__restrict__ t_col, const typename vt<VT>::val_t *__restrict__ t_val,
    int64_t t_blocks) {
    using T = vt<VT>;
    if constexpr (CS > 0) C = CS;
    const bool tails = CS == 0 && LEN && t_slot != nullptr;
    int64_t slot, base, stride;
    int clen, len;
    const int32_t *cp;
    const typename T::val_t *vp;
    if (tails && static_cast<int64_t>(blockIdx.x) < t_blocks) {
        const int lane = threadIdx.x & (kWave - 1);
        const int64_t w = static_cast<int64_t>(blockIdx.x) * (kScsThreads / kWave) +
                          threadIdx.x / kWave;
        const int32_t ts = t_slot[w * kWave + lane];
        if (ts < 0) return;
        slot = ts;
        base = t_ptr[w] + lane;
        stride = kWave;
        clen = chunk_len[ts / C];
        len = slot_len[ts];
        cp = t_col;
        vp = t_val;
    } else {
        int64_t pos, i;
        if (cpb > 0) {
            const int64_t hb = tails ? blockIdx.x - t_blocks : blockIdx.x;
            const int64_t hn = tails ? gridDim.x - t_blocks : gridDim.x;
            const int64_t blk = REMAP ? xcd_block(hb, hn) : hb;
            pos = blk * cpb + threadIdx.x / C;
            i = threadIdx.x % C;
            if (threadIdx.x >= cpb * C) return;
        } else {
            pos = blockIdx.x / bpc;
            i = (blockIdx.x % bpc) * kScsThreads + threadIdx.x;
            if (i >= C) return;
        }
        if (pos >= n_chunks) return;
        const int64_t chunk = order ? static_cast<int64_t>(order[pos]) : pos;
        if (chunk < 0) return;
        slot = chunk * C + i;
        base = static_cast<int64_t>(chunk_ptr[chunk]) + i;
        stride = C;
        clen = chunk_len[chunk];
        len = LEN ? slot_len[slot] : clen;
        if (tails && len > t_head[chunk]) return;
        cp = col;
        vp = val;
    }
    typename T::vec_t acc;
    if constexpr (CS > 0)
        acc = slot_chain_pipe<VT, kScsPipeUnroll<VT>>(cp, vp, x, base, CS, len);
    else
        acc = slot_chain_pipe<VT, kScsUnroll>(cp, vp, x, base, stride, len);
    if (LEN && len < clen) acc = T::fma_(T::widen(typename T::val_t(0)), x[0], acc);
    if (UNPERM) {
        const int32_t r = slot_row[slot];
        if (r < n_rows) y[r] = acc;
    } else {
        y[slot] = acc;
    }
}

// SpMM, X and Y row-major with k columns: thread (slot, v), v fastest, so k
// consecutive lanes share a slot's val/col (one broadcast load) and read a
// contiguous row of X.  Y in original row order.  Workgroups take whole
// chunks (cpb chunks of C*k threads each, or one chunk in bpc parts), in the
// optional execution `order`.  LEN: each slot stops at its row's length and
// no padding term is applied, so Y equals the CRS SpMM bit for bit; without
// it the chain runs over the padded chunk (padding adds 0 * X[0, v]).  The
// chain is software-pipelined like slot_chain_pipe.
template <int VT, bool LEN>
__global__ __launch_bounds__(kScsThreads) void scs_spmm_kernel(
    int64_t C, int64_t n_chunks, int64_t n_rows, int64_t k, int64_t cpb, int64_t bpc,
    const int32_t *__restrict__ chunk_ptr, const int32_t *__restrict__ chunk_len,
    const int32_t *__restrict__ slot_len, const int32_t *__restrict__ order,
    const int32_t *__restrict__ col, const typename vt<VT>::val_t *__restrict__ val,
    const int32_t *__restrict__ slot_row, const typename vt<VT>::vec_t *__restrict__ X,
    typename vt<VT>::vec_t *__restrict__ Y) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    using V = typename T::val_t;
    constexpr int U = 8;
    const int64_t ck = C * k;
    int64_t pos, li;
    if (cpb > 1) {
        if (threadIdx.x >= cpb * ck) return;
        pos = static_cast<int64_t>(blockIdx.x) * cpb + threadIdx.x / ck;
        li = threadIdx.x % ck;
    } else {
        pos = blockIdx.x / bpc;
        li = (blockIdx.x % bpc) * kScsThreads + threadIdx.x;
    }
    if (pos >= n_chunks || li >= ck) return;
    const int64_t chunk = order ? static_cast<int64_t>(order[pos]) : pos;
    const int64_t i = li / k, v = li - i * k;
    const int64_t slot = chunk * C + i;
    const int32_t r = slot_row[slot];
    if (r >= n_rows) return;
    const int64_t base = static_cast<int64_t>(chunk_ptr[chunk]) + i;
    const int len = LEN ? slot_len[slot] : chunk_len[chunk];
    A acc = 0;
    int j = 0;
    if (len >= U) {
        int32_t c0[U];
        V w0[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c0[u] = col[base + u * C];
            w0[u] = val[base + u * C];
        }
        for (; j + 2 * U <= len; j += U) {
            int32_t c1[U];
            V w1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c1[u] = col[base + (j + U + u) * C];
                w1[u] = val[base + (j + U + u) * C];
            }
            A xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = X[static_cast<int64_t>(c0[u]) * k + v];
#pragma unroll
            for (int u = 0; u < U; ++u) acc = T::fma_(T::widen(w0[u]), xv[u], acc);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c0[u] = c1[u];
                w0[u] = w1[u];
            }
        }
        A xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = X[static_cast<int64_t>(c0[u]) * k + v];
#pragma unroll
        for (int u = 0; u < U; ++u) acc = T::fma_(T::widen(w0[u]), xv[u], acc);
        j += U;
    }
    for (; j < len; ++j) {
        const int64_t o = base + static_cast<int64_t>(j) * C;
        acc = T::fma_(T::widen(val[o]), X[static_cast<int64_t>(col[o]) * k + v], acc);
    }
    Y[static_cast<int64_t>(r) * k + v] = acc;
}

// As scs_spmm_kernel for k = K a power of two with 16-byte aligned X and Y:
// a lane owns Q = 16 / sizeof(vec) consecutive columns v (K / Q lanes per
// slot), reads its part of each X row with one 16-byte load and keeps Q fma
// chains, each the same sequential chain as in scs_spmm_kernel.
template <int VT, bool LEN, int K>
__global__ __launch_bounds__(kScsThreads) void scs_spmm_vec(
    int64_t C, int64_t n_chunks, int64_t n_rows, int64_t cpb, int64_t bpc,
    const int32_t *__restrict__ chunk_ptr, const int32_t *__restrict__ chunk_len,
    const int32_t *__restrict__ slot_len, const int32_t *__restrict__ order,
    const int32_t *__restrict__ col, const typename vt<VT>::val_t *__restrict__ val,
    const int32_t *__restrict__ slot_row, const typename vt<VT>::vec_t *__restrict__ X,
    typename vt<VT>::vec_t *__restrict__ Y) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    using V = typename T::val_t;
    constexpr int Q = 16 / sizeof(A);
    constexpr int GL = K / Q;
    constexpr int U = 8;
    static_assert(GL * Q == K, "K must be a multiple of the run length");
    const int64_t ck = C * GL;
    int64_t pos, li;
    if (cpb > 1) {
        if (threadIdx.x >= cpb * ck) return;
        pos = static_cast<int64_t>(blockIdx.x) * cpb + threadIdx.x / ck;
        li = threadIdx.x % ck;
    } else {
        pos = blockIdx.x / bpc;
        li = (blockIdx.x % bpc) * kScsThreads + threadIdx.x;
    }
    if (pos >= n_chunks || li >= ck) return;
    const int64_t chunk = order ? static_cast<int64_t>(order[pos]) : pos;
    const int64_t i = li / GL;
    const int l = static_cast<int>(li % GL);
    const int64_t slot = chunk * C + i;
    const int32_t r = slot_row[slot];
    if (r >= n_rows) return;
    const int64_t base = static_cast<int64_t>(chunk_ptr[chunk]) + i;
    const int len = LEN ? slot_len[slot] : chunk_len[chunk];
    A acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0;
    int j = 0;
    if constexpr ((GL == U || GL == 2 * U) && sizeof(V) <= 4) {
        // K / Q = 8 or 16 lanes per slot (k = 32 or 64 with 4-byte X
        // entries): every lane would load the same column index and value of
        // each element, a third of the bytes the slot requests.  Instead lane
        // l loads element l % 8 of each group of 8 and a swizzle within the 8
        // lanes (a DPP row broadcast within the 16) hands every lane all of
        // them; the next group's are requested before this group's X rows
        // (lomax band N = 1M, k = 32: FP32 values 0.378 -> 0.371 ms, A^T
        // 0.213 -> 0.200 ms; BF16 values 0.389 -> 0.372 ms).
        if (len >= U) {
            const int le = l % U;  // the element of each group this lane loads
            int cm = col[base + static_cast<int64_t>(le) * C];
            int wm = val_bits<V>(val[base + static_cast<int64_t>(le) * C]);
            for (;;) {
                const bool more = j + 2 * U <= len;
                int cn = 0, wn = 0;
                if (more) {
                    const int64_t o = base + static_cast<int64_t>(j + U + le) * C;
                    cn = col[o];
                    wn = val_bits<V>(val[o]);
                }
                int cs[U], ws[U];
                if constexpr (GL == U)
                    grp8_bcast8(cm, wm, cs, ws, std::make_integer_sequence<int, U>{});
                else
                    row16_bcast8(cm, wm, cs, ws, std::make_integer_sequence<int, U>{});
                A xv[U][Q];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    load_run<A, Q>(X + static_cast<int64_t>(cs[u]) * K + l * Q, xv[u]);
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int q = 0; q < Q; ++q)
                        acc[q] = T::fma_(T::widen(bits_val<V>(ws[u])), xv[u][q], acc[q]);
                j += U;
                if (!more) break;
                cm = cn;
                wm = wn;
            }
        }
    } else if (len >= U) {
        int32_t c0[U];
        V w0[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c0[u] = col[base + u * C];
            w0[u] = val[base + u * C];
        }
        for (;;) {
            const bool more = j + 2 * U <= len;
            int32_t c1[U];
            V w1[U];
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    c1[u] = col[base + static_cast<int64_t>(j + U + u) * C];
                    w1[u] = val[base + static_cast<int64_t>(j + U + u) * C];
                }
            }
            A xv[U][Q];
#pragma unroll
            for (int u = 0; u < U; ++u)
                load_run<A, Q>(X + static_cast<int64_t>(c0[u]) * K + l * Q, xv[u]);
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[q] = T::fma_(T::widen(w0[u]), xv[u][q], acc[q]);
            j += U;
            if (!more) break;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c0[u] = c1[u];
                w0[u] = w1[u];
            }
        }
    }
    for (; j < len; ++j) {
        const int64_t o = base + static_cast<int64_t>(j) * C;
        A xv[Q];
        load_run<A, Q>(X + static_cast<int64_t>(col[o]) * K + l * Q, xv);
        const A w = T::widen(val[o]);
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] = T::fma_(w, xv[q], acc[q]);
    }
    store_run<A, Q>(Y + static_cast<int64_t>(r) * K + l * Q, acc);
}

template <int VT, int K>
static void launch_scs_spmm_vec(smax_hip_ctx ctx, int64_t C, int64_t n_chunks, int64_t n_rows,
                                int64_t cpb, int64_t bpc, int64_t blocks, int64_t lds_bytes,
                                const int32_t *chunk_ptr, const int32_t *chunk_len,
                                const int32_t *slot_len, const int32_t *order,
                                const int32_t *col, const typename vt<VT>::val_t *val,
                                const int32_t *slot_row, const typename vt<VT>::vec_t *X,
                                typename vt<VT>::vec_t *Y) {
    const unsigned lds = static_cast<unsigned>(lds_bytes);
    if (slot_len)
        hipLaunchKernelGGL((scs_spmm_vec<VT, true, K>), dim3(static_cast<unsigned>(blocks)),
                           dim3(kScsThreads), lds, ctx->stream, C, n_chunks, n_rows, cpb, bpc,
                           chunk_ptr, chunk_len, slot_len, order, col, val, slot_row, X, Y);
    else
        hipLaunchKernelGGL((scs_spmm_vec<VT, false, K>), dim3(static_cast<unsigned>(blocks)),
                           dim3(kScsThreads), lds, ctx->stream, C, n_chunks, n_rows, cpb, bpc,
                           chunk_ptr, chunk_len, slot_len, order, col, val, slot_row, X, Y);
}

// lds_bytes: dynamic LDS reserved per workgroup of the 16-byte-run kernel
// (an occupancy cap, 0 = none).
template <int VT>
static int launch_scs_spmm(smax_hip_ctx ctx, int64_t C, int64_t n_chunks, int64_t n_rows,
                           int64_t k, const int32_t *chunk_ptr, const int32_t *chunk_len,
                           const int32_t *slot_len, const int32_t *order,
                           const int32_t *col, const void *val, const int32_t *slot_row,
                           const void *X, void *Y, int64_t lds_bytes = 0) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    constexpr int64_t Q = 16 / sizeof(A);
    // k a power of two in [4Q, 64] with 16-byte aligned operands: Q columns
    // per lane, at least 4 lanes per slot
    const bool vec = k >= 4 * Q && k <= 64 && (k & (k - 1)) == 0 && aligned16(X, Y);
    const int64_t ck = C * (vec ? k / Q : k);
    const int64_t cpb = ck <= kScsThreads ? kScsThreads / ck : 1;
    const int64_t bpc = ck <= kScsThreads ? 1 : ceil_div(ck, kScsThreads);
    const int64_t blocks = cpb > 1 ? ceil_div(n_chunks, cpb) : n_chunks * bpc;
    SMAX_HIP_REQUIRE(ctx, blocks <= INT32_MAX, "scs spmm: grid too large");
    auto v = static_cast<const typename T::val_t *>(val);
    auto xv = static_cast<const typename T::vec_t *>(X);
    auto yv = static_cast<typename T::vec_t *>(Y);
    if (vec) {
        switch (k) {
#define SMAX_SCS_SPMM_VEC(KK)                                                                \
    case KK:                                                                                  \
        if constexpr (KK >= Q)                                                                \
            launch_scs_spmm_vec<VT, KK>(ctx, C, n_chunks, n_rows, cpb, bpc, blocks, lds_bytes, \
                                        chunk_ptr, chunk_len, slot_len, order, col, v,         \
                                        slot_row, xv, yv);                                     \
        break;
            SMAX_SCS_SPMM_VEC(2)
            SMAX_SCS_SPMM_VEC(4)
            SMAX_SCS_SPMM_VEC(8)
            SMAX_SCS_SPMM_VEC(16)
            SMAX_SCS_SPMM_VEC(32)
            SMAX_SCS_SPMM_VEC(64)
#undef SMAX_SCS_SPMM_VEC
        }
        SMAX_HIP_LAUNCHED(ctx, "scs_spmm_vec");
        return SMAX_HIP_SUCCESS;
    }
    if (slot_len)
        hipLaunchKernelGGL((scs_spmm_kernel<VT, true>), dim3(static_cast<unsigned>(blocks)),
                           dim3(kScsThreads), 0, ctx->stream, C, n_chunks, n_rows, k, cpb,
                           bpc, chunk_ptr, chunk_len, slot_len, order, col, v, slot_row,
                           xv, yv);
    else
        hipLaunchKernelGGL((scs_spmm_kernel<VT, false>), dim3(static_cast<unsigned>(blocks)),
                           dim3(kScsThreads), 0, ctx->stream, C, n_chunks, n_rows, k, cpb,
                           bpc, chunk_ptr, chunk_len, slot_len, order, col, v, slot_row,
                           xv, yv);
    SMAX_HIP_LAUNCHED(ctx, "scs_spmm_kernel");
    return SMAX_HIP_SUCCESS;
}

struct W64Args {
    int64_t lds_bytes;  // dynamic LDS reserved per workgroup (occupancy cap)
    int64_t n_chunks, n_rows;
    const int32_t *chunk_ptr, *chunk_len, *slot_len, *order, *col;
    const void *val;
    const int32_t *slot_row;
    const void *x;
    void *y;
    const smax_hip_scs_tails *tails;  // nullptr or n_blocks == 0: none
    bool short_chains;                // every chunk at most kShortChain long
};

template <int VT, bool UNPERM, bool REMAP, bool LEN, bool SHORT>
static void launch_w64(smax_hip_ctx ctx, const W64Args &a) {
    using T = vt<VT>;
    const smax_hip_scs_tails *t =
        !SHORT && LEN && a.tails && a.tails->n_blocks > 0 ? a.tails : nullptr;
    const int64_t tb = t ? t->n_blocks : 0;
    hipLaunchKernelGGL((scs_spmv_w64<VT, UNPERM, REMAP, LEN, SHORT>),
                       dim3(static_cast<unsigned>(ceil_div(a.n_chunks, kScsThreads / kWave) + tb)),
                       dim3(kScsThreads), static_cast<unsigned>(a.lds_bytes), ctx->stream,
                       a.n_chunks, a.n_rows, a.chunk_ptr, a.chunk_len, a.slot_len, a.order, a.col,
                       static_cast<const typename T::val_t *>(a.val), a.slot_row,
                       static_cast<const typename T::vec_t *>(a.x),
                       static_cast<typename T::vec_t *>(a.y), t ? t->head : nullptr,
                       t ? t->slot : nullptr, t ? t->ptr : nullptr, t ? t->col : nullptr,
                       static_cast<const typename T::val_t *>(t ? t->val : nullptr), tb);
}

template <int VT, bool UNPERM, bool SHORT>
static void launch_w64_len(smax_hip_ctx ctx, bool remap, const W64Args &a) {
    const bool len = a.slot_len != nullptr;
    if (remap && len)
        launch_w64<VT, UNPERM, true, true, SHORT>(ctx, a);
    else if (remap)
        launch_w64<VT, UNPERM, true, false, SHORT>(ctx, a);
    else if (len)
        launch_w64<VT, UNPERM, false, true, SHORT>(ctx, a);
    else
        launch_w64<VT, UNPERM, false, false, SHORT>(ctx, a);
}

// Short chains only without tail waves (a layout with them is padded more
// than 6x, and its chunks then run their heads only).
template <int VT, bool UNPERM>
static void launch_w64_sched(smax_hip_ctx ctx, bool remap, const W64Args &a) {
    const bool has_tails = a.slot_len && a.tails && a.tails->n_blocks > 0;
    if (a.short_chains && !has_tails)
        launch_w64_len<VT, UNPERM, true>(ctx, remap, a);
    else
        launch_w64_len<VT, UNPERM, false>(ctx, remap, a);
}

// remap: XCD-contiguous order (right when sigma-windows keep rows local);
// slot_len: optional per-slot lengths (nullptr: full chunk chains);
// order: optional execution order for C = 64 (nullptr: chunk order);
// tails: the handle's tail waves (C = 64 with slot_len; nullptr: none);
// max_chunk_len: the longest chunk when known (-1: unknown).
template <int VT>
int launch_scs_spmv(smax_hip_ctx ctx, int64_t C, int64_t n_chunks,
                    int64_t n_rows, const int32_t *chunk_ptr,
                    const int32_t *chunk_len, const int32_t *col,
                    const void *val, const int32_t *slot_row, const void *x,
                    void *y, bool unperm, bool remap = true,
                    const int32_t *slot_len = nullptr,
                    const int32_t *order = nullptr, int64_t lds_bytes = 0,
                    const smax_hip_scs_tails *tails = nullptr, int32_t max_chunk_len = -1) {
    using T = vt<VT>;
    if (C == kWave) {
        const bool short_chains = max_chunk_len >= 0 && max_chunk_len <= kShortChain;
        const W64Args a{lds_bytes, n_chunks, n_rows, chunk_ptr, chunk_len, slot_len,
                        order, col, val, slot_row, x, y, tails, short_chains};
        if (unperm)
            launch_w64_sched<VT, true>(ctx, remap, a);
        else
            launch_w64_sched<VT, false>(ctx, remap, a);
        SMAX_HIP_LAUNCHED(ctx, "scs_spmv_w64");
        return SMAX_HIP_SUCCESS;
    }
    auto v = static_cast<const typename T::val_t *>(val);
    auto xv = static_cast<const typename T::vec_t *>(x);
    auto yv = static_cast<typename T::vec_t *>(y);
    const int64_t cpb = C <= kScsThreads ? kScsThreads / C : 0;
    const int64_t bpc = C <= kScsThreads ? 1 : ceil_div(C, kScsThreads);
    const smax_hip_scs_tails *t =
        slot_len && cpb > 0 && tails && tails->n_blocks > 0 ? tails : nullptr;
    const int64_t tb = t ? t->n_blocks : 0;
    const int64_t blocks = (cpb > 0 ? ceil_div(n_chunks, cpb) : n_chunks * bpc) + tb;
    SMAX_HIP_REQUIRE(ctx, blocks <= INT32_MAX, "scs spmv: grid too large");
    const dim3 grid(static_cast<unsigned>(blocks));
    const bool rm = remap && cpb > 0;
    auto tv = static_cast<const typename T::val_t *>(t ? t->val : nullptr);
    // C = 32 without tail waves: the compile-time chunk width, and the
    // schedule's occupancy cap
    const bool c32 = C == 32 && !t;
    const unsigned lds = c32 ? static_cast<unsigned>(lds_bytes) : 0u;
#define SMAX_SCS_GENERIC_CS(U, L, R, CS)                                                 \
    hipLaunchKernelGGL((scs_spmv_generic<VT, U, L, R, CS>), grid, dim3(kScsThreads), lds, \
                       ctx->stream, C, n_chunks, n_rows, cpb, bpc, chunk_ptr,           \
                       chunk_len, slot_len, order, col, v, slot_row, xv, yv,            \
                       t ? t->head : nullptr, t ? t->slot : nullptr, t ? t->ptr : nullptr, \
                       t ? t->col : nullptr, tv, tb)
#define SMAX_SCS_GENERIC(U, L, R)                                                        \
    if (c32) SMAX_SCS_GENERIC_CS(U, L, R, 32); else SMAX_SCS_GENERIC_CS(U, L, R, 0)
    if (unperm && slot_len) {
        if (rm) SMAX_SCS_GENERIC(true, true, true); else SMAX_SCS_GENERIC(true, true, false);
    } else if (unperm) {
        if (rm) SMAX_SCS_GENERIC(true, false, true); else SMAX_SCS_GENERIC(true, false, false);
    } else if (slot_len) {
        if (rm) SMAX_SCS_GENERIC(false, true, true); else SMAX_SCS_GENERIC(false, true, false);
    } else {
        if (rm) SMAX_SCS_GENERIC(false, false, true); else SMAX_SCS_GENERIC(false, false, false);
    }
#undef SMAX_SCS_GENERIC
#undef SMAX_SCS_GENERIC_CS
    SMAX_HIP_LAUNCHED(ctx, "scs_spmv_generic");
    return SMAX_HIP_SUCCESS;
}

// Execution schedule of the C = 64 kernel on a converted matrix.
//
// Block order: windows much shorter than an XCD's share keep rows (and their
// x bands) local, so each XCD gets a contiguous range of chunks (remap).  A
// near-global sort puts the longest chunks first, where a contiguous split
// would leave one XCD with all of them, and runs each length's rows across
// the whole matrix at once (every XCD then gathers from all of x).  When the
// rows of the slots are known (a converted matrix), such layouts run their
// chunks sorted by their first slot's row instead: the XCDs get contiguous
// ranges of that sequence, so each holds one stretch of rows with its share
// of every length, and rows in flight stay local as with short windows.
// Without the rows, the hardware's round-robin order is kept.
//
// Chunk order: one wave runs a chunk's 64 rows; the longest chunks (rows of
// up to thousands of entries) then set the kernel's critical path, and in
// chunk order the last window's long chunks start late.  Within each range
// (each XCD's, or the whole grid) chunks of at least kLongChunk entries per
// slot (and 8x the mean) start first, longest first, one per workgroup where
// the range has room (each on its own CU, with its own miss queue); the rest
// follow in chunk order.  Results do not depend on the order.
constexpr int kLongChunk = 128;
constexpr int64_t kLdsPerCu = 160 * 1024;

static int upload_order(smax_hip_ctx ctx, const std::vector<int32_t> &ord, int32_t **out) {
    int rc = dev_alloc(ctx, reinterpret_cast<void **>(out), sizeof(int32_t) * ord.size(),
                       "scs execution order");
    if (rc) return rc;
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(*out, ord.data(), sizeof(int32_t) * ord.size(),
                                      hipMemcpyHostToDevice, ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    return SMAX_HIP_SUCCESS;
}

void free_schedule(smax_hip_scs_sched &s) {
    dev_free(s.order);
    dev_free(s.order_lf);
    s.order = s.order_lf = nullptr;
}

int build_schedule(smax_hip_ctx ctx, int64_t C, int64_t sigma, int64_t n_rows_padded,
                   int64_t n_chunks, int64_t n_elements, int64_t nnz,
                   const int32_t *chunk_lengths, smax_hip_scs_sched *out,
                   const int32_t *slot_row) {
    *out = smax_hip_scs_sched{};
    const bool local = sigma * 16 <= n_rows_padded;
    const bool by_row = !local && slot_row != nullptr && C <= kScsThreads;
    out->remap = local || by_row;
    // Occupancy: with little padding every lane of a wave has gathers in
    // flight, and five waves per SIMD congest the path from L2 to the MALL;
    // reserving LDS for four workgroups per CU (four waves per SIMD) runs
    // faster (0.70 -> 0.63 ms on the 4M-row Lomax matrix at sigma = 65536;
    // C = 32: 0.660 -> 0.640 ms).  Padded layouts (fill > 1.5) keep lanes
    // idle and want every wave (C = 32, sigma = 4096, fill 1.7: 0.745 ->
    // 0.824 ms with the cap).
    if ((C == kWave || C == 32) && nnz > 0 && 2 * n_elements <= 3 * nnz)
        out->lds_cap_bytes = kLdsPerCu / 4;
    const int64_t n = n_chunks;
    if (n <= 0) return SMAX_HIP_SUCCESS;
    std::vector<int32_t> cl(static_cast<size_t>(n));
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(cl.data(), chunk_lengths, sizeof(int32_t) * n,
                                      hipMemcpyDeviceToHost, ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    out->max_chunk_len = *std::max_element(cl.begin(), cl.end());
    const double mean =
        static_cast<double>(n_elements) / static_cast<double>(n * C);
    const int64_t thr =
        std::max<int64_t>(kLongChunk, static_cast<int64_t>(std::ceil(8 * mean)));
    const int64_t n_long =
        std::count_if(cl.begin(), cl.end(), [&](int32_t l) { return l >= thr; });
    if (n_long == 0 && !by_row) return SMAX_HIP_SUCCESS;
    // by_row: the chunks sorted by the row of their first slot (stable)
    std::vector<int32_t> seq;
    if (by_row) {
        std::vector<int32_t> first(static_cast<size_t>(n));
        SMAX_HIP_CALL(ctx, hipMemcpy2DAsync(first.data(), sizeof(int32_t), slot_row,
                                            sizeof(int32_t) * C, sizeof(int32_t), n,
                                            hipMemcpyDeviceToHost, ctx->stream));
        SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
        seq.resize(static_cast<size_t>(n));
        for (int64_t c = 0; c < n; ++c) seq[c] = static_cast<int32_t>(c);
        std::stable_sort(seq.begin(), seq.end(),
                         [&](int32_t p, int32_t q) { return first[p] < first[q]; });
    }
    // positions [a, b) of `ord` receive the chunks at [a, min(b, n)) of the
    // chunk sequence (`sq`, or chunk order when null): long chunks first
    // (longest first) at every `spread`-th position when they fit, then the
    // rest in sequence order
    auto fill = [&](std::vector<int32_t> &ord, int64_t a, int64_t b, int64_t spread,
                    const std::vector<int32_t> *sq) {
        const int64_t e = std::min(b, n);
        if (e <= a) return;
        std::vector<int32_t> lg, rest;
        for (int64_t i = a; i < e; ++i) {
            const int32_t c = sq ? (*sq)[i] : static_cast<int32_t>(i);
            (cl[c] >= thr ? lg : rest).push_back(c);
        }
        std::stable_sort(lg.begin(), lg.end(),
                         [&](int32_t p, int32_t q) { return cl[p] > cl[q]; });
        const int64_t m = static_cast<int64_t>(lg.size());
        const int64_t step = m > 0 && (m - 1) * spread < e - a ? spread : 1;
        std::vector<char> taken(static_cast<size_t>(e - a), 0);
        for (int64_t i = 0; i < m; ++i) {
            ord[a + i * step] = lg[i];
            taken[i * step] = 1;
        }
        int64_t r = 0;
        for (int64_t p = a; p < e; ++p)
            if (!taken[p - a]) ord[p] = rest[r++];
    };
    if (n_long > 0) {
        std::vector<int32_t> lf(static_cast<size_t>(n), -1);
        fill(lf, 0, n, 1, nullptr);
        const int rc = upload_order(ctx, lf, &out->order_lf);
        if (rc) return rc;
    }
    if (C > kScsThreads) return SMAX_HIP_SUCCESS;
    const std::vector<int32_t> *sq = by_row ? &seq : nullptr;
    const int64_t per_block = kScsThreads / C;  // chunks per workgroup
    const int64_t nb = ceil_div(n, per_block);
    std::vector<int32_t> ord(static_cast<size_t>(nb * per_block), -1);
    if (out->remap) {
        constexpr int64_t kXcd = 8;
        const int64_t q = nb / kXcd, rem = nb % kXcd;
        for (int64_t k = 0; k < kXcd; ++k) {
            const int64_t s = k < rem ? k * (q + 1) : rem * (q + 1) + (k - rem) * q;
            const int64_t len = k < rem ? q + 1 : q;
            fill(ord, s * per_block, (s + len) * per_block, per_block, sq);
        }
    } else {
        fill(ord, 0, nb * per_block, per_block, sq);
    }
    return upload_order(ctx, ord, &out->order);
}

// ------------------------------------------------------------ tail waves
// Layouts whose padding exceeds 5x their stored entries get tail waves (see
// smax_hip_scs_tails).  Measured on the 4M-row band matrix (FP64): SELL-64-1
// 2.32 -> 1.14 ms, SELL-64-64 2.09 -> 1.04 ms, SELL-32-1 2.14 -> 1.17 ms,
// SELL-32-64 1.91 -> 1.05 ms, SELL-128-1 2.20 -> 1.13 ms.  Sorted windows
// leave only the matrix's longest rows skewed: their waves then carry the
// kernel's longest chains together and finish last (SELL-64-4096, fill 2.5:
// 0.76 ms against 0.70; SELL-128-4096, fill 4.0: 0.82 against 0.75).  A
// skewed chunk keeps its slots up to its (C/8)-th longest one (the 8th for
// C = 64: 1.10 ms for SELL-64-1, against 1.23 ms up to the 4th);
// a tail wave takes at most kTailWaveWork elements (32768: 1.10 ms, 16384:
// 1.11 ms, unbounded 1.39 ms).
constexpr int64_t kTailWaveWork = 32768;

// Copies the elements of every tail slot into its wave's arrays: element j of
// the slot (chunk c, lane i) at chunk_ptr[c] + j*C + i goes to
// t_ptr[w] + j*64 + l for its tail-wave lane l of wave w.  One thread per
// tail-wave lane; t_col null: values only.
template <typename E>
__global__ void k_tail_fill(int64_t n_lanes, int64_t C, const int32_t *__restrict__ t_slot,
                            const int64_t *__restrict__ t_ptr,
                            const int32_t *__restrict__ slot_len,
                            const int32_t *__restrict__ chunk_ptr,
                            const int32_t *__restrict__ col, const E *__restrict__ val,
                            int32_t *__restrict__ t_col, E *__restrict__ t_val) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n_lanes) return;
    const int32_t ts = t_slot[p];
    if (ts < 0) return;
    const int64_t src = static_cast<int64_t>(chunk_ptr[ts / C]) + ts % C;
    const int64_t dst = t_ptr[p / kWave] + p % kWave;
    const int len = slot_len[ts];
    for (int j = 0; j < len; ++j) {
        const int64_t s = static_cast<int64_t>(j) * C, d = static_cast<int64_t>(j) * kWave;
        if (t_col) t_col[dst + d] = col[src + s];
        t_val[dst + d] = val[src + s];
    }
}

static int tail_fill(smax_hip_ctx ctx, smax_hip_scs A, bool with_col) {
    const smax_hip_scs_info &in = A->info;
    const smax_hip_scs_tails &t = A->tails;
    const int64_t lanes = t.n_blocks * kScsThreads;
    int32_t *tc = with_col ? t.col : nullptr;
#define SMAX_TAIL_FILL(E)                                                                 \
    hipLaunchKernelGGL((k_tail_fill<E>), grid_for(lanes, kScsThreads), dim3(kScsThreads), \
                       0, ctx->stream, lanes, in.C, t.slot, t.ptr, in.slot_len,           \
                       in.chunk_ptr,                                                      \
                       in.col, static_cast<const E *>(in.val), tc, static_cast<E *>(t.val))
    switch (in.val_bytes) {
    case 8: SMAX_TAIL_FILL(uint64_t); break;
    case 4: SMAX_TAIL_FILL(uint32_t); break;
    default: SMAX_TAIL_FILL(uint16_t); break;
    }
#undef SMAX_TAIL_FILL
    SMAX_HIP_LAUNCHED(ctx, "k_tail_fill");
    return SMAX_HIP_SUCCESS;
}

void free_tails(smax_hip_scs_tails &t) {
    dev_free(t.head);
    dev_free(t.slot);
    dev_free(t.ptr);
    dev_free(t.col);
    dev_free(t.val);
    t = smax_hip_scs_tails{};
}

int build_tails(smax_hip_ctx ctx, smax_hip_scs A) {
    const smax_hip_scs_info &in = A->info;
    const int64_t C = in.C;
    if (C < 16 || C > kScsThreads || C % 8 != 0 || in.nnz <= 0 || in.n_elements <= 6 * in.nnz)
        return SMAX_HIP_SUCCESS;
    const int64_t rank = C / 8;
    const int64_t nc = in.n_chunks;
    std::vector<int32_t> sl(static_cast<size_t>(nc * C)), cl(static_cast<size_t>(nc));
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(sl.data(), in.slot_len, sizeof(int32_t) * sl.size(),
                                      hipMemcpyDeviceToHost, ctx->stream));
    SMAX_HIP_CALL(ctx, hipMemcpyAsync(cl.data(), in.chunk_lengths, sizeof(int32_t) * nc,
                                      hipMemcpyDeviceToHost, ctx->stream));
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    // per chunk: the head length, and the slots beyond it
    std::vector<int32_t> head(static_cast<size_t>(nc));
    std::vector<std::pair<int32_t, int32_t>> tail;  // (length, slot)
    std::vector<int32_t> lens(static_cast<size_t>(C));
    for (int64_t c = 0; c < nc; ++c) {
        const int32_t *l = sl.data() + c * C;
        std::copy(l, l + C, lens.begin());
        std::nth_element(lens.begin(), lens.begin() + (rank - 1), lens.end(),
                         std::greater<int32_t>());
        const int32_t h = lens[rank - 1];
        head[c] = 2 * static_cast<int64_t>(h) < cl[c] ? h : cl[c];
        for (int64_t i = 0; i < C; ++i)
            if (l[i] > head[c]) tail.emplace_back(l[i], static_cast<int32_t>(c * C + i));
    }
    if (tail.empty()) return SMAX_HIP_SUCCESS;
    // longest first (ties in slot order), packed greedily into waves
    std::sort(tail.begin(), tail.end(), [](const auto &a, const auto &b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    std::vector<int64_t> wave_len;  // each wave's longest slot (its first)
    std::vector<int64_t> pos(tail.size());
    int64_t cnt = kWave, work = 0;
    for (size_t t = 0; t < tail.size(); ++t) {
        const int64_t l = tail[t].first;
        if (cnt == kWave || work + l > kTailWaveWork) {
            wave_len.push_back(l);
            cnt = work = 0;
        }
        pos[t] = static_cast<int64_t>(wave_len.size() - 1) * kWave + cnt;
        ++cnt;
        work += l;
    }
    const int64_t n_waves = static_cast<int64_t>(wave_len.size());
    const int64_t per_block = kScsThreads / kWave;
    const int64_t n_blocks = ceil_div(ceil_div(n_waves, per_block), 8) * 8;
    std::vector<int32_t> tslot(static_cast<size_t>(n_blocks * kScsThreads), -1);
    for (size_t t = 0; t < tail.size(); ++t) tslot[pos[t]] = tail[t].second;
    std::vector<int64_t> tptr(static_cast<size_t>(n_blocks * per_block), 0);
    int64_t store = 0;
    for (int64_t w = 0; w < n_waves; ++w) {
        tptr[w] = store;
        store += wave_len[w] * kWave;
    }
    smax_hip_scs_tails t;
    t.n_elements = store;
    t.n_slots = static_cast<int64_t>(tail.size());
    t.n_waves = n_waves;
    int rc = 0;
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&t.head), sizeof(int32_t) * nc, "scs tails");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&t.slot), sizeof(int32_t) * tslot.size(),
                    "scs tails");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&t.ptr), sizeof(int64_t) * tptr.size(),
                    "scs tails");
    rc |= dev_alloc(ctx, reinterpret_cast<void **>(&t.col), sizeof(int32_t) * store, "scs tails");
    rc |= dev_alloc(ctx, &t.val, static_cast<size_t>(in.val_bytes) * store, "scs tails");
    if (rc) {
        free_tails(t);
        return SMAX_HIP_ERROR_OUT_OF_MEMORY;
    }
    // elements past a lane's length are never read; zero them all the same
    rc = smax_hip_zero(ctx, t.col, sizeof(int32_t) * store);
    if (!rc) rc = smax_hip_zero(ctx, t.val, static_cast<size_t>(in.val_bytes) * store);
    if (rc) {
        free_tails(t);
        return rc;
    }
    const hipError_t e1 = hipMemcpyAsync(t.head, head.data(), sizeof(int32_t) * nc,
                                         hipMemcpyHostToDevice, ctx->stream);
    const hipError_t e2 = hipMemcpyAsync(t.slot, tslot.data(), sizeof(int32_t) * tslot.size(),
                                         hipMemcpyHostToDevice, ctx->stream);
    const hipError_t e3 = hipMemcpyAsync(t.ptr, tptr.data(), sizeof(int64_t) * tptr.size(),
                                         hipMemcpyHostToDevice, ctx->stream);
    const hipError_t e4 = hipStreamSynchronize(ctx->stream);
    for (hipError_t e : {e1, e2, e3, e4})
        if (e != hipSuccess) {
            free_tails(t);
            return hip_fail(ctx, e, "scs tails");
        }
    t.n_blocks = n_blocks;
    A->tails = t;  // freed with the handle from here on
    rc = tail_fill(ctx, A, true);
    if (rc) return rc;
    SMAX_HIP_CALL(ctx, hipStreamSynchronize(ctx->stream));
    return SMAX_HIP_SUCCESS;
}

int refresh_tail_values(smax_hip_ctx ctx, smax_hip_scs A) {
    return A->tails.n_blocks > 0 ? tail_fill(ctx, A, false) : SMAX_HIP_SUCCESS;
}

static int check_scs_args(smax_hip_ctx ctx, const char *fn, int value_type,
                          int64_t C, int64_t n_chunks, const void *chunk_ptr,
                          const void *chunk_len, const void *col,
                          const void *val, const void *x, const void *y) {
    if (!valid_value_type(value_type))
        return fail(ctx, SMAX_HIP_ERROR_UNSUPPORTED, "%s: value type %d", fn,
                    value_type);
    SMAX_HIP_REQUIRE(ctx, C > 0 && n_chunks > 0, "%s: bad C/n_chunks", fn);
    SMAX_HIP_REQUIRE(ctx, C <= INT32_MAX && n_chunks <= INT32_MAX / C, "%s: too many slots", fn);
    SMAX_HIP_REQUIRE(ctx, chunk_ptr && chunk_len && col && val && x && y,
                     "%s: null pointer", fn);
    return SMAX_HIP_SUCCESS;
}

}  // namespace smaxhip

using namespace smaxhip;

extern "C" {

int smax_hip_scs_spmv(smax_hip_ctx ctx, int value_type, int64_t C,
                      int64_t n_chunks, const int32_t *chunk_ptr,
                      const int32_t *chunk_lengths, const int32_t *col,
                      const void *val, const void *x, void *y) {
    SMAX_HIP_CHECK_CTX(ctx);
    int rc = check_scs_args(ctx, "smax_hip_scs_spmv", value_type, C, n_chunks,
                            chunk_ptr, chunk_lengths, col, val, x, y);
    if (rc) return rc;
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_scs_spmv<VT>(ctx, C, n_chunks,
                                                    n_chunks * C, chunk_ptr,
                                                    chunk_lengths, col, val,
                                                    nullptr, x, y, false));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_scs_spmv_unpermuted(smax_hip_ctx ctx, int value_type, int64_t C,
                                 int64_t n_chunks, int64_t n_rows,
                                 const int32_t *chunk_ptr,
                                 const int32_t *chunk_lengths,
                                 const int32_t *col, const void *val,
                                 const int32_t *slot_row, const void *x,
                                 void *y) {
    SMAX_HIP_CHECK_CTX(ctx);
    int rc = check_scs_args(ctx, "smax_hip_scs_spmv_unpermuted", value_type, C,
                            n_chunks, chunk_ptr, chunk_lengths, col, val, x, y);
    if (rc) return rc;
    SMAX_HIP_REQUIRE(ctx, slot_row && n_rows > 0 && n_rows <= n_chunks * C,
                     "smax_hip_scs_spmv_unpermuted: bad slot_row / n_rows");
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_scs_spmv<VT>(ctx, C, n_chunks, n_rows,
                                                    chunk_ptr, chunk_lengths,
                                                    col, val, slot_row, x, y,
                                                    true));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_scs_plan_create(smax_hip_ctx ctx, int64_t C, int64_t sigma,
                             int64_t n_rows_padded, int64_t n_chunks,
                             int64_t n_elements, int64_t nnz,
                             const int32_t *chunk_lengths, smax_hip_scs_plan *out) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, out, "smax_hip_scs_plan_create: null output");
    *out = nullptr;
    SMAX_HIP_REQUIRE(ctx, C > 0 && sigma > 0 && n_chunks > 0 && n_elements >= 0 &&
                              nnz >= 0 && nnz <= n_elements && C <= INT32_MAX &&
                              n_chunks <= INT32_MAX / C && n_rows_padded == n_chunks * C &&
                              chunk_lengths,
                     "smax_hip_scs_plan_create: bad arguments");
    auto *plan = new smax_hip_scs_plan_s();
    plan->C = C;
    plan->n_chunks = n_chunks;
    int rc = build_schedule(ctx, C, sigma, n_rows_padded, n_chunks, n_elements, nnz,
                            chunk_lengths, &plan->sched);
    if (rc) {
        free_schedule(plan->sched);
        delete plan;
        return rc;
    }
    *out = plan;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_scs_plan_destroy(smax_hip_ctx ctx, smax_hip_scs_plan plan) {
    if (!plan) return SMAX_HIP_SUCCESS;
    if (ctx) {
        (void)bind(ctx);
        (void)hipStreamSynchronize(ctx->stream);
    }
    free_schedule(plan->sched);
    delete plan;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_scs_spmv_plan(smax_hip_ctx ctx, int value_type, smax_hip_scs_plan plan,
                           const int32_t *chunk_ptr, const int32_t *chunk_lengths,
                           const int32_t *col, const void *val, const void *x,
                           void *y) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, plan, "smax_hip_scs_spmv_plan: null plan");
    int rc = check_scs_args(ctx, "smax_hip_scs_spmv_plan", value_type, plan->C,
                            plan->n_chunks, chunk_ptr, chunk_lengths, col, val, x, y);
    if (rc) return rc;
    const int64_t npad = plan->n_chunks * plan->C;
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_scs_spmv<VT>(ctx, plan->C, plan->n_chunks, npad,
                                                    chunk_ptr, chunk_lengths, col, val,
                                                    nullptr, x, y, false,
                                                    plan->sched.remap, nullptr,
                                                    plan->C <= kScsThreads
                                                        ? plan->sched.order
                                                        : plan->sched.order_lf,
                                                    plan->sched.lds_cap_bytes, nullptr,
                                                    plan->sched.max_chunk_len));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_scs_apply(smax_hip_ctx ctx, int value_type, smax_hip_scs A,
                       const void *x, void *y, int unpermuted) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, A, "smax_hip_scs_apply: null matrix");
    const smax_hip_scs_info &in = A->info;
    int rc = check_scs_args(ctx, "smax_hip_scs_apply", value_type, in.C, in.n_chunks,
                            in.chunk_ptr, in.chunk_lengths, in.col, in.val, x, y);
    if (rc) return rc;
    if (vt_val_bytes(value_type) != in.val_bytes)
        return fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_scs_apply: value type %d does not match %lld-byte values",
                    value_type, (long long)in.val_bytes);
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_scs_spmv<VT>(ctx, in.C, in.n_chunks, in.n_rows,
                                                    in.chunk_ptr, in.chunk_lengths,
                                                    in.col, in.val, in.slot_row, x, y,
                                                    unpermuted != 0, A->sched.remap,
                                                    in.slot_len,
                                                    in.C <= kScsThreads
                                                        ? A->sched.order
                                                        : A->sched.order_lf,
                                                    A->sched.lds_cap_bytes, &A->tails,
                                                    A->sched.max_chunk_len));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_scs_get_tails(smax_hip_scs A, int64_t *n_slots, int64_t *n_waves,
                           int64_t *n_entries) {
    if (!A || !n_slots || !n_waves || !n_entries)
        return fail(nullptr, SMAX_HIP_ERROR_INVALID_VALUE, "smax_hip_scs_get_tails: null");
    *n_slots = A->tails.n_slots;
    *n_waves = A->tails.n_waves;
    *n_entries = A->tails.n_elements;
    return SMAX_HIP_SUCCESS;
}

int smax_hip_scs_spmm(smax_hip_ctx ctx, int value_type, int64_t C,
                      int64_t n_chunks, int64_t n_rows,
                      const int32_t *chunk_ptr, const int32_t *chunk_lengths,
                      const int32_t *col, const void *val,
                      const int32_t *slot_row, int64_t k, const void *X,
                      void *Y) {
    SMAX_HIP_CHECK_CTX(ctx);
    int rc = check_scs_args(ctx, "smax_hip_scs_spmm", value_type, C, n_chunks,
                            chunk_ptr, chunk_lengths, col, val, X, Y);
    if (rc) return rc;
    SMAX_HIP_REQUIRE(ctx, slot_row && k > 0 && n_rows > 0 &&
                              n_rows <= n_chunks * C,
                     "smax_hip_scs_spmm: bad slot_row / n_rows / k");
    if (!extent_ok(n_chunks * C, k))
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "smax_hip_scs_spmm: k = %lld exceeds the launch limits", (long long)k);
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_scs_spmm<VT>(ctx, C, n_chunks, n_rows, k, chunk_ptr,
                                                    chunk_lengths, nullptr, nullptr, col,
                                                    val, slot_row, X, Y));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

int smax_hip_scs_apply_spmm(smax_hip_ctx ctx, int value_type, smax_hip_scs A,
                            int64_t k, const void *X, void *Y) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, A, "smax_hip_scs_apply_spmm: null matrix");
    SMAX_HIP_REQUIRE(ctx, !A->from_sell,
                     "smax_hip_scs_apply_spmm: the matrix was built from SELL arrays, whose slot "
                     "lengths may drop stored +0 entries; build it with smax_hip_scs_from_crs");
    const smax_hip_scs_info &in = A->info;
    int rc = check_scs_args(ctx, "smax_hip_scs_apply_spmm", value_type, in.C,
                            in.n_chunks, in.chunk_ptr, in.chunk_lengths, in.col, in.val,
                            X, Y);
    if (rc) return rc;
    SMAX_HIP_REQUIRE(ctx, k > 0, "smax_hip_scs_apply_spmm: k must be positive");
    if (!extent_ok(in.n_rows_padded, k))
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "smax_hip_scs_apply_spmm: k = %lld exceeds the launch limits", (long long)k);
    if (vt_val_bytes(value_type) != in.val_bytes)
        return fail(ctx, SMAX_HIP_ERROR_INVALID_VALUE,
                    "smax_hip_scs_apply_spmm: value type %d does not match %lld-byte "
                    "values", value_type, (long long)in.val_bytes);
    // the SpMV's occupancy cap pays off for the SpMM too when long chunks lead
    // (0.40 -> 0.37 ms on the 1M-row band matrix at sigma = 65536, k = 32);
    // without long chunks it costs (0.27 -> 0.28 ms on its transpose)
    const int64_t lds = A->sched.order_lf ? A->sched.lds_cap_bytes : 0;
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_scs_spmm<VT>(ctx, in.C, in.n_chunks, in.n_rows, k,
                                                    in.chunk_ptr, in.chunk_lengths,
                                                    in.slot_len, A->sched.order_lf, in.col,
                                                    in.val, in.slot_row, X, Y, lds));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

}  // extern "C"
