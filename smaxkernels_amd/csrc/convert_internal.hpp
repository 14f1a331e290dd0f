// Declarations shared by the conversion, SpMV, adjoint and solver units.
#pragma once

#include "common.hpp"

struct smax_hip_scs_s {
    smax_hip_scs_info info;
    // Built on first use by the CG solvers, freed with the handle:
    int32_t *col_slot = nullptr;  // perm[col[e]]: columns in slot space
    float *val_f32 = nullptr;     // values rounded to binary32
    uint16_t *val_bf16 = nullptr; // values rounded to bfloat16
    float *diag_f32 = nullptr;    // diagonal per slot, binary32
    int bf16_split = -1;          // 1: val_bf16 holds 0 on the diagonal
    // Execution order of the C = 64 SpMV (block slot -> chunk, -1 = none;
    // nullptr = chunk order) and whether it assumes the XCD-contiguous block
    // remap.  Built with the conversion.
    int32_t *order = nullptr;
    bool order_remap = false;
    // Chunk order of the SpMM kernels (any C): long chunks first, longest
    // first, then the rest in chunk order; nullptr = chunk order.
    int32_t *order_lf = nullptr;
};

namespace smaxhip {

// CRS -> SELL-C-sigma.  Element j of the CRS arrays takes its value (and its
// src_index) from position src_map[j] of `val` when src_map is non-null,
// from position j otherwise.
int build_scs(smax_hip_ctx ctx, int64_t n_rows, int64_t n_cols, int64_t nnz,
              const int32_t *row_ptr, const int32_t *col, const void *val,
              int64_t val_bytes, const int32_t *src_map, int64_t C,
              int64_t sigma, smax_hip_scs *out);

void scs_free_arrays(smax_hip_scs_info &info);

// Builds scs->order / scs->order_remap from the chunk lengths (synchronous).
int build_chunk_order(smax_hip_ctx ctx, smax_hip_scs scs);

}  // namespace smaxhip
