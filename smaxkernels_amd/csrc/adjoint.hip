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
// element e of SELL row r; a group of up to 64 lanes takes one element, each
// lane a strided part of v (fma chain), then a fixed xor-butterfly sums the
// group: the reduction tree depends only on k.  Workgroups take contiguous
// element ranges, so the chunk of an element is found by stepping, not by a
// search per element.
#include <utility>

#include "common.hpp"
#include "convert_internal.hpp"
#include "radix.hpp"
#include "scan.hpp"

namespace smaxhip {

// ------------------------------------------------------------------ SDDMM
// Workgroup b takes elements [b*seg, (b+1)*seg) in storage order; its groups
// step through them, tracking the chunk of the current element (one binary
// search over chunk_ptr per thread at the start, then forward steps).
template <int VT, int G>
__global__ __launch_bounds__(256) void scs_sddmm_kernel(
    int64_t n_elements, int64_t C, int64_t n_chunks, int64_t k, int64_t seg,
    const int32_t *__restrict__ chunk_ptr, const int32_t *__restrict__ col,
    const int32_t *__restrict__ slot_row, const int32_t *__restrict__ src,
    const typename vt<VT>::vec_t *__restrict__ Gm,
    const typename vt<VT>::vec_t *__restrict__ X,
    typename vt<VT>::vec_t *__restrict__ dvals) {
    using T = vt<VT>;
    using A = typename T::vec_t;
    constexpr int kGroups = 256 / G;
    const int64_t e0 = static_cast<int64_t>(blockIdx.x) * seg;
    const int64_t e1 = e0 + seg < n_elements ? e0 + seg : n_elements;
    const int group = threadIdx.x / G;
    const int lane = threadIdx.x % G;
    // chunk a with chunk_ptr[a] <= e0 < chunk_ptr[a+1] (empty chunks skipped)
    int64_t a = 0, b = n_chunks;
    while (b - a > 1) {
        const int64_t m = (a + b) >> 1;
        if (chunk_ptr[m] <= e0) a = m; else b = m;
    }
    int64_t next = chunk_ptr[a + 1];
    for (int64_t e = e0 + group; e < e1; e += kGroups) {
        while (next <= e) next = chunk_ptr[++a + 1];
        const int32_t s = src[e];
        A acc = 0;
        if (s >= 0) {
            const int64_t off = e - chunk_ptr[a];
            const int64_t slot = a * C + (C == kWave ? (off & (kWave - 1)) : off % C);
            const int64_t r = slot_row[slot];
            const int64_t c = col[e];
            for (int64_t v = lane; v < k; v += G)
                acc = T::fma_(Gm[r * k + v], X[c * k + v], acc);
        }
#pragma unroll
        for (int off = G / 2; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, G);
        if (s >= 0 && lane == 0) dvals[s] = acc;
    }
}

template <int VT, int G>
static void launch_sddmm_g(smax_hip_ctx ctx, const smax_hip_scs_info &in,
                           int64_t k, const void *Gm, const void *X, void *dv) {
    using T = vt<VT>;
    const int64_t seg = (256 / G) * 16;
    hipLaunchKernelGGL((scs_sddmm_kernel<VT, G>), grid_for(in.n_elements, seg),
                       dim3(256), 0, ctx->stream, in.n_elements, in.C,
                       in.n_chunks, k, seg, in.chunk_ptr, in.col, in.slot_row,
                       in.src_index, static_cast<const typename T::vec_t *>(Gm),
                       static_cast<const typename T::vec_t *>(X),
                       static_cast<typename T::vec_t *>(dv));
}

template <int VT>
static int launch_sddmm(smax_hip_ctx ctx, const smax_hip_scs_info &in, int64_t k,
                        int g, const void *Gm, const void *X, void *dv) {
    switch (g) {
    case 1: launch_sddmm_g<VT, 1>(ctx, in, k, Gm, X, dv); break;
    case 2: launch_sddmm_g<VT, 2>(ctx, in, k, Gm, X, dv); break;
    case 4: launch_sddmm_g<VT, 4>(ctx, in, k, Gm, X, dv); break;
    case 8: launch_sddmm_g<VT, 8>(ctx, in, k, Gm, X, dv); break;
    case 16: launch_sddmm_g<VT, 16>(ctx, in, k, Gm, X, dv); break;
    case 32: launch_sddmm_g<VT, 32>(ctx, in, k, Gm, X, dv); break;
    default: launch_sddmm_g<VT, 64>(ctx, in, k, Gm, X, dv); break;
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
    const smax_hip_scs_info &in = A->info;
    if (in.n_elements == 0) return SMAX_HIP_SUCCESS;
    int g = 1;
    while (g < k && g < 64) g <<= 1;
    SMAX_HIP_DISPATCH_VT(value_type,
                         return launch_sddmm<VT>(ctx, in, k, g, G, X, dvals));
    return SMAX_HIP_ERROR_UNSUPPORTED;
}

}  // extern "C"
