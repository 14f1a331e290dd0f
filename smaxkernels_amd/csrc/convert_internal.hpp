// Declarations shared by the conversion, SpMV, adjoint and solver units.
#pragma once

#include "common.hpp"

// Execution schedule of the SELL kernels, built once from the chunk lengths.
struct smax_hip_scs_sched {
    // C = 64 SpMV: wave position -> chunk (-1 = none), nullptr = chunk order;
    // `remap` says whether it assumes the XCD-contiguous block remap.
    int32_t *order = nullptr;
    bool remap = false;
    // SpMM and C != 64 SpMV: long chunks first, longest first, then the rest
    // in chunk order; nullptr = chunk order.
    int32_t *order_lf = nullptr;
    // Dynamic LDS reserved by the C = 64 (and C = 32) SpMV launch to cap its
    // occupancy (0 = no cap).
    int64_t lds_cap_bytes = 0;
    // Longest chunk (-1: unknown).  Layouts whose chunks all fit one short
    // group run the C = 64 SpMV's single-group chains.
    int32_t max_chunk_len = -1;
};

// Tail waves of a layout whose chunks mix rows of very different lengths
// (built with the conversion for 16 <= C <= 256, C a multiple of 8, when
// padding exceeds 5x the stored entries).  In a chunk whose (C/8)-th longest
// slot is under half the chunk's length, the slots longer than that one are
// taken out of the chunk and repacked, longest first, into waves of their
// own: a SELL-64 whose element j of lane l of wave w sits at
// ptr[w] + j*64 + l, a wave taking slots while it has fewer than 64 and their
// lengths sum to at most kTailWaveWork.  The chunk then runs only as long as
// its (C/8)-th longest slot.  Each slot's chain is still one lane's
// sequential chain over the same elements in the same order.
struct smax_hip_scs_tails {
    int32_t *head = nullptr;  // per chunk: longer slots run in tail waves
    int32_t *slot = nullptr;  // per tail-wave lane (n_blocks * 256): slot, -1 none
    int64_t *ptr = nullptr;   // per tail wave: its first element
    int32_t *col = nullptr;   // the tail waves' elements
    void *val = nullptr;
    int64_t n_blocks = 0;     // workgroups of tail waves (a multiple of 8; 0: none)
    int64_t n_elements = 0;   // storage of col / val
    int64_t n_slots = 0, n_waves = 0;
};

struct smax_hip_scs_s {
    smax_hip_scs_info info;
    // Built on first use by the CG solvers, freed with the handle:
    int32_t *col_slot = nullptr;  // perm[col[e]]: columns in slot space
    float *val_f32 = nullptr;     // values rounded to binary32
    uint16_t *val_bf16 = nullptr; // values rounded to bfloat16
    float *diag_f32 = nullptr;    // diagonal per slot, binary32
    int bf16_split = -1;          // 1: val_bf16 holds 0 on the diagonal
    // Built on first use by the CG solvers when every stored element's slot-
    // space column lies within +-32767 of its slot: col_slot - slot as int16
    // (0 for padding).  col_rel_state: 0 not tried, 1 built, -1 does not fit.
    int16_t *col_rel = nullptr;
    int col_rel_state = 0;
    // Built after col_rel when C = 64 and every chunk's elements use at most
    // 32 distinct offsets: per chunk a table of its offsets (32 int16 entries)
    // and per element a one-byte index into its chunk's table.  The indices
    // of elements j..j+3 (j a multiple of 4) of one slot share a 32-bit word:
    // chunk c's words start at word row col_idx_row[c] (rows of 64 words, one
    // per slot) and element (j, lane) is byte j % 4 of word (col_idx_row[c] +
    // j / 4) * 64 + lane.  col_tab_state: 0 not tried, 1 built, -1 does not
    // fit.
    uint32_t *col_idx = nullptr;
    int64_t *col_idx_row = nullptr;
    int64_t col_idx_rows = 0;     // word rows in all
    int16_t *col_tab = nullptr;
    int col_tab_state = 0;
    smax_hip_scs_sched sched;     // built with the conversion
    // src_index of a slot's stored elements is consecutive (the conversion
    // took values in CRS order, without a source map): element j of a slot
    // whose row starts at CRS position p has src_index p + j.
    bool src_consecutive = false;
    // Built on first use by the SDDMM: row of every stored element (-1 for
    // padding), so an element's row needs no search.
    int32_t *elem_row = nullptr;
    smax_hip_scs_tails tails;     // built with the conversion
    // Built by smax_hip_scs_from_sell from caller SELL arrays: no src_index
    // (no CRS positions), slot lengths derived from the arrays.  Only the SpMV
    // runs on such a handle.
    bool from_sell = false;
};

struct smax_hip_scs_plan_s {
    smax_hip_scs_sched sched;
    int64_t C = 0, n_chunks = 0;
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

// Builds the execution schedule of a SELL layout from its device chunk
// lengths (synchronous); free_schedule releases it.  slot_row (device, may be
// null) gives each slot's original row, which lets layouts sorted over wide
// windows run their chunks in row order.
int build_schedule(smax_hip_ctx ctx, int64_t C, int64_t sigma, int64_t n_rows_padded,
                   int64_t n_chunks, int64_t n_elements, int64_t nnz,
                   const int32_t *chunk_lengths, smax_hip_scs_sched *out,
                   const int32_t *slot_row = nullptr);
void free_schedule(smax_hip_scs_sched &s);

// Builds A's tail waves when its layout qualifies (synchronous); copies the
// values into them again after a value refresh (asynchronous); frees them.
int build_tails(smax_hip_ctx ctx, smax_hip_scs A);
int refresh_tail_values(smax_hip_ctx ctx, smax_hip_scs A);
void free_tails(smax_hip_scs_tails &t);

}  // namespace smaxhip
