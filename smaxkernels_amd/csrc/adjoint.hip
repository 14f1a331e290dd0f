// Differentiation support: the SELL-C-sigma layout of A^T and the SDDMM
// value gradient on A's SELL layout.  Neither has a reference counterpart.
//
// A^T: the CRS positions j = 0..nnz-1 are already in row-major order, so a
// *stable* sort of (col[j], j) by column yields A^T's CRS with every column's
// entries in ascending original row order.  The sort is an LSD radix sort
// with 4-bit digits whose scatter ranks come from prefix sums only (no
// atomics decide an order), so the layout is deterministic.  The resulting
// CRS of A^T goes through the same GPU SELL conversion as A, with src_index
// pointing back at A's CRS nonzeros.
//
// SDDMM: dvals[src_index[e]] = sum_v G[r*k+v] * X[col[e]*k+v] for every stored
// element e of SELL row r.  The reduction tree depends only on k: with P the
// smallest power of two >= k, g = clamp(P / 4, 1, 64) lanes share an element,
// lane l sums v in [l*Q, min(k, l*Q + Q)), Q = ceil(k / g), with an fma chain
// from zero, and an xor-butterfly adds the g partial sums.  For k a power of
// two up to 64 the lanes of a group share one row (SpMM order): the G row is
// read once per row and each element costs one coalesced X row with 16-byte
// loads.  Other k take one group per stored element, with the row of each
// element from a per-element row array built on first use.  Both kernels
// evaluate the same tree, so dvals do not depend on which one runs.
#include <algorithm>
#include <utility>

#include "common.hpp"
#include "convert_internal.hpp"
#include "radix.hpp"
#include "scan.hpp"

namespace smaxhip {

// ------------------------------------------------------------------ SDDMM
// elem_row[e] = original row of stored element e, -1 for padding; one lane
// per slot walks its chunk column (stores coalesced across the chunk).
__global__ __launch_bounds__(256) void k_elem_row(
    int64_t C, int64_t n_slots, int64_t n_rows, const int32_t *__restrict__ chunk_ptr,
    const int32_t *__restrict__ chunk_len, const int32_t *__restrict__ slot_len,
    const int32_t *__restrict__ slot_row, int32_t *__restrict__ elem_row) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (s >= n_slots) return;
    const int64_t c = s / C, i = s - c * C;
    const int64_t base = static_cast<int64_t>(chunk_ptr[c]) + i;
    const int32_t r = slot_row[s];
    const int len = r < n_rows ? slot_len[s] : 0;
    const int clen = chunk_len[c];
    for (int j = 0; j < clen; ++j) elem_row[base + static_cast<int64_t>(j) * C] = j < len ? r : -1;
}

// Lanes per element and per-lane run length of the SDDMM tree for k.
__host__ __device__ constexpr int64_t sddmm_group(int64_t k) {
    int64_t p = 1;
    while (p < k) p <<= 1;
    return p / 4 < 1 ? 1 : (p / 4 > 64 ? 64 : p / 4);
}
__host__ __device__ constexpr int64_t sddmm_run(int64_t k) {
    return (k + sddmm_group(k) - 1) / sddmm_group(k);
}

// The xor-butterfly over G lanes: every lane ends with the group's sum.
template <int G, typename A> __device__ __forceinline__ A group_sum(A acc) {
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, G);
    return acc;
}

// One group of G = sddmm_group(k) lanes per stored element.
template <int VT, int G>
__global__ __launch_bounds__(256) void scs_sddmm_kernel(
    int64_t n_elements, int64_t k, const int32_t *__restrict__ col,
    const int32_t *__restrict__ elem_row, const int32_t *__restrict__ src,
    const typename vt<VT>::vec_t *__restrict__ Gm,
    const typename vt<VT>::vec_t *__restrict__ X,
    typename vt<VT>::vec_t *__restrict__ dvals) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t e = t / G;
    const int lane = static_cast<int>(t % G);
    const bool valid = e < n_elements;
    const int32_t s = valid ? src[e] : -1;
    A acc = 0;
    if (s >= 0) {
        const int64_t r = elem_row[e];
        const int64_t c = col[e];
        const int64_t q = sddmm_run(k);
        const int64_t v1 = (lane + 1) * q < k ? (lane + 1) * q : k;
        for (int64_t v = lane * q; v < v1; ++v) acc = T::fma_(Gm[r * k + v], X[c * k + v], acc);
    }
    acc = group_sum<G>(acc);
    if (s >= 0 && lane == 0) dvals[s] = acc;
}

// k = K a power of two <= 64: GL = sddmm_group(K) lanes per slot (row) in
// SpMM order, long chunks first.  Lane l keeps G[r, l*Q .. l*Q+Q) in
// registers for the whole row; each element costs one coalesced X row read
// with 16-byte loads, Q fmas per lane and the butterfly over GL lanes.  The
// chain over the row's elements is software-pipelined (the next U column and
// src indices are loaded before this group's X rows; U = kSddmmChainGroup),
// and the U results of a
// group, consecutive CRS positions of one row, leave in one store
// instruction from U lanes (U / GL instructions when GL < U).  The U
// elements of a group share one butterfly (group_reduce_scatter; per GL
// elements when GL < U) that ends with each element's sum on the lane that
// stores it, and each lane loads only the src indices it stores to.
// The same butterfly over GL >= U lanes for U elements at once: at each
// offset (GL/2 first) a lane keeps the half of its live elements that its
// lane bit selects and sends the other half to its partner, adding (own +
// partner) as group_sum does, so every element's sum is the same tree.  After
// log2(U) halving steps the remaining offsets (below GL/U) add as in
// group_sum.  Lane l returns the sum of element l / (GL / U).
// The halves are vector values, not array elements: a choice between two
// elements of one array becomes an array read at a lane-dependent index,
// which the compiler expands into a select over every element.
template <int N, int OFF, int GL, typename A>
__device__ __forceinline__ A scatter_steps(const A (&cur)[N], int l) {
    if constexpr (N == 1) {
        return cur[0];
    } else {
        constexpr int H = N / 2;
        const bool upper = (l & OFF) != 0;
        A next[H];
        if constexpr (H == 1) {
            const A keep = upper ? cur[1] : cur[0];
            const A send = upper ? cur[0] : cur[1];
            next[0] = keep + __shfl_xor(send, OFF, GL);
        } else {
            typedef A half_t __attribute__((ext_vector_type(H)));
            half_t lo, hi;
#pragma unroll
            for (int i = 0; i < H; ++i) {
                lo[i] = cur[i];
                hi[i] = cur[H + i];
            }
            const half_t keep = upper ? hi : lo;
            const half_t send = upper ? lo : hi;
#pragma unroll
            for (int i = 0; i < H; ++i) next[i] = keep[i] + __shfl_xor(send[i], OFF, GL);
        }
        return scatter_steps<H, OFF / 2, GL>(next, l);
    }
}

template <int GL, int U, typename A>
__device__ __forceinline__ A group_reduce_scatter(A (&cur)[U], int l) {
    static_assert(GL >= U, "one element per lane at most");
    A s = scatter_steps<U, GL / 2, GL>(cur, l);
#pragma unroll
    for (int off = GL / U / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, GL);
    return s;
}

// Elements per pipelined group of the row kernel, measured at N = 1M (band
// matrix, SELL-64-65536): 16 for k = 16 and 32 (FP64 k = 16 0.61 -> 0.51 ms,
// k = 32 0.80 -> 0.69 ms; FP32 k = 32 0.403 -> 0.385 ms), 8 for k = 64 (FP32
// 0.71 ms against 0.77 with 16) and for k <= 8, where a lane holds several
// elements of a group and 16 would leave three or four waves per SIMD.
template <int K>
constexpr int kSddmmChainGroup = K == 16 || K == 32 ? 16 : 8;

// SEQ: the handle's src_index is consecutive within a slot, so a slot's
// destinations are its first element's src_index plus j (one src load per
// slot instead of one per element).
template <int VT, int K, bool SEQ>
__global__ __launch_bounds__(256) void scs_sddmm_rows(
    int64_t C, int64_t n_chunks, int64_t n_rows, int64_t cpb, int64_t bpc,
    const int32_t *__restrict__ chunk_ptr, const int32_t *__restrict__ slot_len,
    const int32_t *__restrict__ order, const int32_t *__restrict__ col,
    const int32_t *__restrict__ slot_row, const int32_t *__restrict__ src,
    const typename vt<VT>::vec_t *__restrict__ Gm,
    const typename vt<VT>::vec_t *__restrict__ X,
    typename vt<VT>::vec_t *__restrict__ dvals) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    constexpr int GL = static_cast<int>(sddmm_group(K));
    constexpr int Q = K / GL;
    constexpr int U = kSddmmChainGroup<K>;
    static_assert(GL * Q == K, "K must be a power of two");
    const int64_t ck = C * GL;
    int64_t pos, li;
    if (cpb > 1) {
        if (threadIdx.x >= cpb * ck) return;
        pos = static_cast<int64_t>(blockIdx.x) * cpb + threadIdx.x / ck;
        li = threadIdx.x % ck;
    } else {
        pos = blockIdx.x / bpc;
        li = (blockIdx.x % bpc) * 256 + threadIdx.x;
    }
    if (pos >= n_chunks || li >= ck) return;
    const int64_t chunk = order ? static_cast<int64_t>(order[pos]) : pos;
    const int64_t i = li / GL;
    const int l = static_cast<int>(li % GL);
    const int64_t slot = chunk * C + i;
    const int32_t r = slot_row[slot];
    if (r >= n_rows) return;
    const int len = slot_len[slot];
    const int64_t base = static_cast<int64_t>(chunk_ptr[chunk]) + i;
    // SEQ: CRS position of the slot's element 0 (its row start)
    const int64_t row0 = SEQ && len > 0 ? src[base] : 0;
    A g[Q];
    load_run<A, Q>(Gm + static_cast<int64_t>(r) * K + l * Q, g);
    int j = 0;
    if (len >= U) {
        // kPer lanes hold each result, each lane holds kNpl results: lane l
        // stores elements n*GL + l/kPer (n < kNpl) of each group of U
        constexpr int kPer = GL >= U ? GL / U : 1;
        constexpr int kNpl = GL >= U ? 1 : U / GL;
        const int lane_elem = l / kPer;
        const bool stores = l % kPer == 0;
        int32_t c0[U], s0[kNpl];
#pragma unroll
        for (int u = 0; u < U; ++u) c0[u] = col[base + u * C];
#pragma unroll
        for (int n = 0; n < kNpl; ++n)
            s0[n] = SEQ ? 0 : src[base + static_cast<int64_t>(n * GL + lane_elem) * C];
        for (;;) {
            const bool more = j + 2 * U <= len;
            int32_t c1[U], s1[kNpl];
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u)
                    c1[u] = col[base + static_cast<int64_t>(j + U + u) * C];
#pragma unroll
                for (int n = 0; n < kNpl; ++n)
                    s1[n] = SEQ ? 0
                                : src[base + static_cast<int64_t>(j + U + n * GL + lane_elem) * C];
            }
            A xv[U][Q];
#pragma unroll
            for (int u = 0; u < U; ++u)
                load_run<A, Q>(X + static_cast<int64_t>(c0[u]) * K + l * Q, xv[u]);
            A part[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                A acc = 0;
#pragma unroll
                for (int t = 0; t < Q; ++t) acc = T::fma_(g[t], xv[u][t], acc);
                part[u] = acc;
            }
            if constexpr (GL >= U) {
                const A sum = group_reduce_scatter<GL, U>(part, l);
                if (stores) dvals[SEQ ? row0 + j + lane_elem : s0[0]] = sum;
            } else {
#pragma unroll
                for (int n = 0; n < kNpl; ++n) {
                    A piece[GL];
#pragma unroll
                    for (int i = 0; i < GL; ++i) piece[i] = part[n * GL + i];
                    dvals[SEQ ? row0 + j + n * GL + lane_elem : s0[n]] =
                        group_reduce_scatter<GL, GL>(piece, l);
                }
            }
            j += U;
            if (!more) break;
#pragma unroll
            for (int u = 0; u < U; ++u) c0[u] = c1[u];
            for (int n = 0; n < kNpl; ++n) s0[n] = s1[n];
        }
    }
    for (; j < len; ++j) {
        const int64_t e = base + static_cast<int64_t>(j) * C;
        A xv[Q];
        load_run<A, Q>(X + static_cast<int64_t>(col[e]) * K + l * Q, xv);
        A acc = 0;
#pragma unroll
        for (int t = 0; t < Q; ++t) acc = T::fma_(g[t], xv[t], acc);
        acc = group_sum<GL>(acc);
        if (l == 0) dvals[SEQ ? row0 + j : src[e]] = acc;
    }
}

template <int VT, int K>
static void launch_sddmm_rows(smax_hip_ctx ctx, smax_hip_scs A, const void *Gm,
                              const void *X, void *dv) {
    using T = vt<VT>;
    const smax_hip_scs_info &in = A->info;
    const int64_t ck = in.C * sddmm_group(K);
    const int64_t cpb = ck <= 256 ? 256 / ck : 1;
    const int64_t bpc = ck <= 256 ? 1 : ceil_div(ck, 256);
    const int64_t blocks = cpb > 1 ? ceil_div(in.n_chunks, cpb) : in.n_chunks * bpc;
    // the SpMV's occupancy cap, as for the SpMM, when long chunks lead
    const unsigned lds =
        static_cast<unsigned>(A->sched.order_lf ? A->sched.lds_cap_bytes : 0);
#define SMAX_SDDMM_ROWS(SEQ)                                                               \
    hipLaunchKernelGGL((scs_sddmm_rows<VT, K, SEQ>), dim3(static_cast<unsigned>(blocks)), \
                       dim3(256), lds, ctx->stream, in.C, in.n_chunks, in.n_rows, cpb, bpc, \
                       in.chunk_ptr, in.slot_len, A->sched.order_lf, in.col, in.slot_row,  \
                       in.src_index, static_cast<const typename T::vec_t *>(Gm),           \
                       static_cast<const typename T::vec_t *>(X),                          \
                       static_cast<typename T::vec_t *>(dv))
    if (A->src_consecutive)
        SMAX_SDDMM_ROWS(true);
    else
        SMAX_SDDMM_ROWS(false);
#undef SMAX_SDDMM_ROWS
}

template <int VT>
static bool sddmm_rows(smax_hip_ctx ctx, smax_hip_scs A, int64_t k, const void *Gm,
                       const void *X, void *dv) {
    switch (k) {
    case 1: launch_sddmm_rows<VT, 1>(ctx, A, Gm, X, dv); return true;
    case 2: launch_sddmm_rows<VT, 2>(ctx, A, Gm, X, dv); return true;
    case 4: launch_sddmm_rows<VT, 4>(ctx, A, Gm, X, dv); return true;
    case 8: launch_sddmm_rows<VT, 8>(ctx, A, Gm, X, dv); return true;
    case 16: launch_sddmm_rows<VT, 16>(ctx, A, Gm, X, dv); return true;
    case 32: launch_sddmm_rows<VT, 32>(ctx, A, Gm, X, dv); return true;
    case 64: launch_sddmm_rows<VT, 64>(ctx, A, Gm, X, dv); return true;
    default: return false;
    }
}

template <int VT, int G>
static void launch_sddmm_g(smax_hip_ctx ctx, const smax_hip_scs_info &in,
                           const int32_t *elem_row, int64_t k, const void *Gm,
                           const void *X, void *dv) {
    using T = vt<VT>;
    hipLaunchKernelGGL((scs_sddmm_kernel<VT, G>), grid_for(in.n_elements * G, 256),
                       dim3(256), 0, ctx->stream, in.n_elements, k, in.col, elem_row,
                       in.src_index, static_cast<const typename T::vec_t *>(Gm),
                       static_cast<const typename T::vec_t *>(X),
                       static_cast<typename T::vec_t *>(dv));
}

template <int VT>
static int launch_sddmm(smax_hip_ctx ctx, const smax_hip_scs_info &in,
                        const int32_t *elem_row, int64_t k, int g, const void *Gm,
                        const void *X, void *dv) {
    switch (g) {
    case 1: launch_sddmm_g<VT, 1>(ctx, in, elem_row, k, Gm, X, dv); break;
    case 2: launch_sddmm_g<VT, 2>(ctx, in, elem_row, k, Gm, X, dv); break;
    case 4: launch_sddmm_g<VT, 4>(ctx, in, elem_row, k, Gm, X, dv); break;
    case 8: launch_sddmm_g<VT, 8>(ctx, in, elem_row, k, Gm, X, dv); break;
    case 16: launch_sddmm_g<VT, 16>(ctx, in, elem_row, k, Gm, X, dv); break;
    case 32: launch_sddmm_g<VT, 32>(ctx, in, elem_row, k, Gm, X, dv); break;
    default: launch_sddmm_g<VT, 64>(ctx, in, elem_row, k, Gm, X, dv); break;
    }
    SMAX_HIP_LAUNCHED(ctx, "scs_sddmm_kernel");
    return SMAX_HIP_SUCCESS;
}

}  // namespace smaxhip

using namespace smaxhip;

extern "C" {

int smax_hip_scs_transpose_from_crs(smax_hip_ctx ctx, int64_t n_rows,
                                    int64_t n_cols, int64_t nnz,
                                    const int32_t *row_ptr, const int32_t *col,
                                    const void *val, int64_t val_bytes,
                                    int64_t C, int64_t sigma,
                                    smax_hip_scs *out) {
    SMAX_HIP_CHECK_CTX(ctx);
    SMAX_HIP_REQUIRE(ctx, out, "smax_hip_scs_transpose_from_crs: null output");
    *out = nullptr;
    SMAX_HIP_REQUIRE(ctx, n_rows > 0 && n_cols > 0 && nnz >= 0 &&
                              nnz <= INT32_MAX && row_ptr &&
                              (nnz == 0 || (col && val)),
                     "smax_hip_scs_transpose_from_crs: bad input");
    if (n_rows > INT32_MAX || n_cols > INT32_MAX)
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "smax_hip_scs_transpose_from_crs: %lld x %lld exceeds int32",
                    (long long)n_rows, (long long)n_cols);
    const int64_t m = nnz > 0 ? nnz : 1;
    uint32_t *keys = nullptr;
    int32_t *vals = nullptr, *t_row_ptr = nullptr, *t_col = nullptr;
    RadixSort sorter(ctx);
    int rc = sorter.reserve(m);
    if (!rc)
        rc = dev_alloc(ctx, reinterpret_cast<void **>(&t_row_ptr),
                       4 * (n_cols + 1), "transpose");
    if (!rc) rc = dev_alloc(ctx, reinterpret_cast<void **>(&t_col), 4 * m, "transpose");
    if (!rc && nnz > 0) {
        hipLaunchKernelGGL(iota_keys, grid_for(nnz, 256), dim3(256), 0,
                           ctx->stream, nnz, col, sorter.keys(), sorter.vals());
        rc = check_launch(ctx, "iota_keys");
        if (!rc) rc = sorter.sort(nnz, n_cols);
    }
    keys = sorter.keys();
    vals = sorter.vals();
    if (!rc) {
        hipLaunchKernelGGL(transpose_row_ptr, grid_for(n_cols + 1, 256), dim3(256),
                           0, ctx->stream, n_cols, nnz, keys, t_row_ptr);
        rc = check_launch(ctx, "transpose_row_ptr");
    }
    if (!rc && nnz > 0) {
        hipLaunchKernelGGL(transpose_cols, grid_for(nnz, 256), dim3(256), 0,
                           ctx->stream, nnz, n_rows, row_ptr, vals, t_col);
        rc = check_launch(ctx, "transpose_cols");
    }
    if (!rc)
        rc = build_scs(ctx, n_cols, n_rows, nnz, t_row_ptr, t_col, val,
                       val_bytes, vals, C, sigma, out);
    (void)hipStreamSynchronize(ctx->stream);
    dev_free(t_row_ptr);
    dev_free(t_col);
    return rc;
}

int smax_hip_scs_sddmm(smax_hip_ctx ctx, int value_type, smax_hip_scs A,
                       int64_t k, const void *G, const void *X, void *dvals) {
    SMAX_HIP_CHECK_CTX(ctx);
    if (!valid_value_type(value_type))
        return fail(ctx, SMAX_HIP_ERROR_UNSUPPORTED,
                    "smax_hip_scs_sddmm: value type %d", value_type);
    SMAX_HIP_REQUIRE(ctx, A && k > 0 && G && X && dvals,
                     "smax_hip_scs_sddmm: bad arguments");
    SMAX_HIP_REQUIRE(ctx, !A->from_sell,
                     "smax_hip_scs_sddmm: the matrix was built from SELL arrays and has no CRS "
                     "positions; build it with smax_hip_scs_from_crs");
    const smax_hip_scs_info &in = A->info;
    if (!extent_ok(in.n_rows_padded, k) || !extent_ok(std::max<int64_t>(in.n_elements, 1), k))
        return fail(ctx, SMAX_HIP_ERROR_OVERFLOW,
                    "smax_hip_scs_sddmm: k = %lld exceeds the launch limits", (long long)k);
    if (in.n_elements == 0) return SMAX_HIP_SUCCESS;
    // the row kernel's 16-byte loads need 16-byte aligned G and X rows
    const int64_t run_bytes = sddmm_run(k) * (value_type == SMAX_HIP_F64 ? 8 : 4);
    const bool aligned = run_bytes % 16 != 0 || aligned16(G, X);
    if (aligned && (k & (k - 1)) == 0 && k <= 64 && in.C * sddmm_group(k) <= INT32_MAX) {
        bool done = false;
        SMAX_HIP_DISPATCH_VT(value_type, done = sddmm_rows<VT>(ctx, A, k, G, X, dvals));
        if (done) {
            SMAX_HIP_LAUNCHED(ctx, "scs_sddmm_rows");
            return SMAX_HIP_SUCCESS;
        }
    }
    if (!A->elem_row) {
        int rc = dev_alloc(ctx, reinterpret_cast<void **>(&A->elem_row),
                           sizeof(int32_t) * in.n_elements, "sddmm element rows");
        if (rc) return rc;
        hipLaunchKernelGGL(k_elem_row, grid_for(in.n_rows_padded, 256), dim3(256), 0,
                           ctx->stream, in.C, in.n_rows_padded, in.n_rows, in.chunk_ptr,
                           in.chunk_lengths, in.slot_len, in.slot_row, A->elem_row);
        SMAX_HIP_LAUNCHED(ctx, "k_elem_row");
    }
    const int g = static_cast<int>(sddmm_group(k));
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_sddmm<VT>(ctx, in, A->elem_row, k, g, G, X,
                                                 dvals));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

}  // extern "C"
