/*
 * smax_hip.h — C ABI of libsmax_hip.so, the MI355X (gfx950) HIP platform of
 * the SmaxKernels SpMV path.
 *
 * Plain C: opaque handles, raw device pointers, 64-bit sizes, int status
 * returns.  Nothing here throws or exits.  All matrix index arrays are int32
 * (the other SMAX::IntType values are rejected with SMAX_HIP_ERROR_UNSUPPORTED
 * by smax_hip_check_types).  Every compute entry point enqueues work on the
 * context's current stream and returns without synchronising; a non-zero
 * status from one that launches a single kernel means nothing was enqueued,
 * and smax_hip_last_error() describes the failure.  Conversions, plans and
 * solvers allocate and synchronise.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the SmaxKernels source tree).  The reference CUDA platform does
 * allocation + copy in initialize, launch + device-wide synchronise in apply
 * and copy-back in finalize (src/kernels/spmv/spmv_cuda/spmv_cuda_core.cuh:
 * 11-205); the C++ HIP platform layered on this ABI keeps that split, with
 * buffers owned and freed by the platform.
 */
#ifndef SMAX_HIP_H
#define SMAX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMAX_HIP_ABI_VERSION 2

/* ------------------------------------------------------------------ status */
typedef enum {
    SMAX_HIP_SUCCESS = 0,
    SMAX_HIP_ERROR_INVALID_VALUE = 1, /* bad size, null pointer, bad flag   */
    SMAX_HIP_ERROR_UNSUPPORTED = 2,   /* index/value type not built for GPU */
    SMAX_HIP_ERROR_OUT_OF_MEMORY = 3,
    SMAX_HIP_ERROR_LAUNCH = 4,        /* kernel launch rejected             */
    SMAX_HIP_ERROR_RUNTIME = 5,       /* any other HIP runtime error        */
    SMAX_HIP_ERROR_NO_DEVICE = 6,
    SMAX_HIP_ERROR_OVERFLOW = 7,      /* result does not fit int32 indexing */
    SMAX_HIP_ERROR_NOT_CONVERGED = 8  /* solver hit its iteration limit     */
} smax_hip_status;

/*
 * Size limits.  Row, column and slot counts are int32; a product over k
 * right-hand sides (SpMM, SDDMM) also needs k <= INT32_MAX and
 * rows * k <= 2^40 (one thread per output in a grid of 2^32 workgroups).
 * Larger sizes return SMAX_HIP_ERROR_OVERFLOW before anything is read.
 */

/*
 * Value types.  Replaces SMAX::FloatType {FLOAT32, FLOAT64}
 * (src/common.hpp:28) and adds BF16 value storage.
 *   F64 : f64 values, f64 x/y, f64 accumulation
 *   F32 : f32 values, f32 x/y, f32 accumulation
 *   BF16: bf16 values (uint16 bit patterns), f32 x/y, f32 accumulation
 */
typedef enum {
    SMAX_HIP_F64 = 0,
    SMAX_HIP_F32 = 1,
    SMAX_HIP_BF16 = 2
} smax_hip_value_type;

/* Index types, in the order of SMAX::IntType (src/common.hpp:25). */
typedef enum {
    SMAX_HIP_INT16 = 0,
    SMAX_HIP_INT32 = 1,
    SMAX_HIP_INT64 = 2,
    SMAX_HIP_UINT16 = 3,
    SMAX_HIP_UINT32 = 4,
    SMAX_HIP_UINT64 = 5
} smax_hip_index_type;

/* Kernel variants for CRS SpMV.  Both compute every row as a sequential fma
 * chain in CRS order, i.e. bit-identical to the reference CPU row sum. */
typedef enum {
    SMAX_HIP_CRS_SEQUENTIAL = 0, /* lane per row, straight from HBM          */
    SMAX_HIP_CRS_BLOCKED = 1     /* rows binned by length into row blocks:
                                    each workgroup streams one block's
                                    values/columns coalesced through LDS;
                                    short rows are reduced lane per row,
                                    rows above the bin limit stream alone
                                    (workgroup per row)                     */
} smax_hip_crs_kernel;

typedef struct smax_hip_ctx_s *smax_hip_ctx;
typedef struct smax_hip_crs_plan_s *smax_hip_crs_plan;
typedef struct smax_hip_scs_s *smax_hip_scs;
typedef struct smax_hip_scs_plan_s *smax_hip_scs_plan;
typedef struct smax_hip_bcrs_s *smax_hip_bcrs;

/* ---------------------------------------------------------------- context */
int smax_hip_abi_version(void);

/* SMAX_HIP_SUCCESS if (index_type, value_type) runs on the GPU.  Replaces
 * the type switch of dispatch_cuda (src/kernels/spmv/spmv_cuda.hpp:76-134). */
int smax_hip_check_types(int index_type, int value_type);

/* One context per (device, stream).  Replaces the implicit CUDA runtime
 * state used by src/platforms/cuda/cuda.cuh:12-82. */
int smax_hip_ctx_create(int device, smax_hip_ctx *out);
int smax_hip_ctx_destroy(smax_hip_ctx ctx);

/* Message of the last failure on this context (ctx may be NULL: then the
 * calling thread's last failure anywhere in the library). */
const char *smax_hip_last_error(smax_hip_ctx ctx);

/* Run subsequent work on `stream` (a hipStream_t); NULL restores the
 * context's own stream. */
int smax_hip_set_stream(smax_hip_ctx ctx, void *stream);
void *smax_hip_get_stream(smax_hip_ctx ctx);
int smax_hip_sync(smax_hip_ctx ctx);

/* -------------------------------------------------------------- memory */
/* Replace transfer_HtoD / transfer_DtoH (src/platforms/cuda/cuda.cuh:12-82);
 * copies are complete when they return. */
int smax_hip_malloc(smax_hip_ctx ctx, void **dptr, size_t bytes);
int smax_hip_free(smax_hip_ctx ctx, void *dptr);
int smax_hip_memcpy_h2d(smax_hip_ctx ctx, void *dst, const void *src,
                        size_t bytes);
int smax_hip_memcpy_d2h(smax_hip_ctx ctx, void *dst, const void *src,
                        size_t bytes);
int smax_hip_memcpy_d2d(smax_hip_ctx ctx, void *dst, const void *src,
                        size_t bytes);
/* Fills `bytes` bytes with zero using a kernel (safe under graph capture). */
int smax_hip_zero(smax_hip_ctx ctx, void *dptr, size_t bytes);

/* Stream-ordered timing: elapsed milliseconds between a start and a stop
 * recorded on the context's current stream (stop waits for completion). */
int smax_hip_timer_start(smax_hip_ctx ctx);
int smax_hip_timer_stop(smax_hip_ctx ctx, float *elapsed_ms);

/* --------------------------------------------------------------- CRS SpMV */
/*
 * y[r] = sum_j val[j] * x[col[j]] for r < n_rows; overwrites y.
 * Replaces naive_crs_spmv_cuda_launcher
 * (src/kernels/spmv/spmv_cuda/spmv_cuda_crs_impl.cu:24-46) and restates
 * naive_crs_spmv (src/kernels/spmv/spmv_cpu/spmv_cpu_crs_impl.hpp:9-38).
 * plan == NULL runs SMAX_HIP_CRS_SEQUENTIAL, a plan SMAX_HIP_CRS_BLOCKED.
 */
int smax_hip_crs_spmv(smax_hip_ctx ctx, int value_type, int64_t n_rows,
                      int64_t n_cols, int64_t nnz, const int32_t *row_ptr,
                      const int32_t *col, const void *val, const void *x,
                      void *y, smax_hip_crs_plan plan);

/* Plan of one CRS pattern (device row_ptr), built once and reused by every
 * smax_hip_crs_spmv (row blocks, SMAX_HIP_CRS_BLOCKED) and smax_hip_crs_spmm
 * (row execution order) on that pattern.  No reference counterpart: the
 * reference CUDA kernel is thread-per-row only. */
int smax_hip_crs_plan_create(smax_hip_ctx ctx, int64_t n_rows, int64_t nnz,
                             const int32_t *row_ptr, smax_hip_crs_plan *out);
int smax_hip_crs_plan_destroy(smax_hip_ctx ctx, smax_hip_crs_plan plan);

/* --------------------------------------------------------------- CRS SpMM */
/*
 * Y = A X for k right-hand sides.  row_major: X[c*ldx + v], Y[r*ldy + v];
 * column-major: X[v*ldx + c], Y[v*ldy + r].  The reference column-major
 * kernel uses ldx = ldy = n_rows (spmm_cpu_crs_impl.hpp:29,41).
 * Replaces naive_crs_spmm_{col,row}_maj
 * (src/kernels/spmm/spmm_cpu/spmm_cpu_crs_impl.hpp:9-89), which has no CUDA
 * counterpart (src/kernels/spmm.hpp:75-84).  Each output is a sequential fma
 * chain in CRS order (bit-identical to the reference CPU kernel).  A plan of
 * the same pattern (smax_hip_crs_plan_create; may be NULL) makes row-major
 * products run long rows first and the rest by length inside 65536-row
 * windows; results are the same with or without it.  Column-major products
 * with k >= 2 copy X to row-major, run the row-major kernel (with the plan)
 * and copy Y back, in a workspace the context allocates on first use and
 * grows when a larger product needs it (those calls allocate; the others only
 * launch kernels).  An outgrown workspace is kept until the context is
 * destroyed, so work already queued and HIP graphs captured before the growth
 * stay valid; calls that overlap in time on different streams of one context
 * share the workspace and race.  X must hold (k-1)*ldx + n_cols entries
 * (row-major: (n_cols-1)*ldx + k).  Leading dimensions: row-major ldx, ldy
 * >= k; column-major ldx >= min(n_rows, n_cols) (columns of X may overlap,
 * as with the reference's ldx = n_rows when n_cols > n_rows) and, for
 * k >= 2, ldy >= n_rows (columns of Y must not overlap).
 */
int smax_hip_crs_spmm(smax_hip_ctx ctx, int value_type, int64_t n_rows,
                      int64_t n_cols, int64_t nnz, const int32_t *row_ptr,
                      const int32_t *col, const void *val, int64_t k,
                      int row_major, const void *X, int64_t ldx, void *Y,
                      int64_t ldy, smax_hip_crs_plan plan);

/*
 * smax_hip_crs_spmm with no allocation of any kind, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Same arguments, checks and results (bit for bit) as smax_hip_crs_spmm.
 * Column-major products with k >= 2 stage X and Y through a fixed 32 MiB
 * device array of the library (one per device, held until the process
 * ends) instead of the context's workspace, in passes of as many columns as
 * fit: (n_rows + n_cols) * sizeof(vector type) * columns <= 32 MiB - 256.
 * A product whose single column does not fit (n_rows + n_cols > 4,194,272
 * for F64, 8,388,544 for F32/BF16) is refused with
 * SMAX_HIP_ERROR_INVALID_VALUE.  Every call on one device shares that array,
 * whatever its context: calls that overlap in time on one device (other
 * streams, other contexts, other threads) race and must be serialised by the
 * caller.  Row-major products and k = 1 use no scratch.
 */
int smax_hip_crs_spmm_noalloc(smax_hip_ctx ctx, int value_type, int64_t n_rows,
                              int64_t n_cols, int64_t nnz, const int32_t *row_ptr,
                              const int32_t *col, const void *val, int64_t k,
                              int row_major, const void *X, int64_t ldx, void *Y,
                              int64_t ldy, smax_hip_crs_plan plan);

/* ------------------------------------------------------------ SELL-C-sigma */
/*
 * y[c*C + i] = sum_{j < chunk_lengths[c]} val[o] * x[col[o]],
 * o = chunk_ptr[c] + j*C + i, for every chunk c and lane i < C: y holds
 * n_chunks*C entries in SELL (permuted, padded) order.
 * Replaces naive_scs_spmv_cuda_launcher
 * (src/kernels/spmv/spmv_cuda/spmv_cuda_scs_impl.cu:30-54) and restates
 * naive_scs_spmv (src/kernels/spmv/spmv_cpu/spmv_cpu_scs_impl.hpp:9-49).
 * Each lane is a sequential fma chain (bit-identical to the reference).
 */
int smax_hip_scs_spmv(smax_hip_ctx ctx, int value_type, int64_t C,
                      int64_t n_chunks, const int32_t *chunk_ptr,
                      const int32_t *chunk_lengths, const int32_t *col,
                      const void *val, const void *x, void *y);

/*
 * As smax_hip_scs_spmv, but y is written in the original row order:
 * y[slot_row[s]] for every slot s with slot_row[s] < n_rows (slot_row maps a
 * SELL slot to its original row, i.e. the inverse of perm; padding slots
 * hold values >= n_rows).  No reference counterpart (the reference returns
 * permuted y, examples/validation/mkl/mkl_scs_spmv_validate.cpp:87-89).
 */
int smax_hip_scs_spmv_unpermuted(smax_hip_ctx ctx, int value_type, int64_t C,
                                 int64_t n_chunks, int64_t n_rows,
                                 const int32_t *chunk_ptr,
                                 const int32_t *chunk_lengths,
                                 const int32_t *col, const void *val,
                                 const int32_t *slot_row, const void *x,
                                 void *y);

/*
 * Execution schedule for smax_hip_scs_spmv_plan on caller-owned SELL arrays
 * (the arrays the reference registers, spmv.hpp:37-50), built once from the
 * device chunk_lengths: long chunks start first, for sigma*16 <=
 * n_rows_padded each XCD takes a contiguous range of chunks, and the fill
 * n_elements / nnz decides the occupancy of the C = 64 kernel.  The plan holds
 * only the schedule, so it stays valid while the arrays keep their chunk
 * structure (new values are fine).  No reference counterpart.
 */
int smax_hip_scs_plan_create(smax_hip_ctx ctx, int64_t C, int64_t sigma,
                             int64_t n_rows_padded, int64_t n_chunks,
                             int64_t n_elements, int64_t nnz,
                             const int32_t *chunk_lengths, smax_hip_scs_plan *out);
int smax_hip_scs_plan_destroy(smax_hip_ctx ctx, smax_hip_scs_plan plan);

/* smax_hip_scs_spmv through a plan: same output, bit for bit. */
int smax_hip_scs_spmv_plan(smax_hip_ctx ctx, int value_type, smax_hip_scs_plan plan,
                           const int32_t *chunk_ptr, const int32_t *chunk_lengths,
                           const int32_t *col, const void *val, const void *x,
                           void *y);

/*
 * Y = A X on a SELL matrix, X (n_cols x k) and Y (n_rows x k) row-major, Y
 * in original row order through slot_row.  Each output is a sequential fma
 * chain over the row's padded SELL elements (a padding element adds
 * 0 * X[0, v]).  No reference counterpart; see smax_hip_scs_apply_spmm for
 * the chain over stored entries only.
 */
int smax_hip_scs_spmm(smax_hip_ctx ctx, int value_type, int64_t C,
                      int64_t n_chunks, int64_t n_rows,
                      const int32_t *chunk_ptr, const int32_t *chunk_lengths,
                      const int32_t *col, const void *val,
                      const int32_t *slot_row, int64_t k, const void *X,
                      void *Y);

/*
 * Y = A X on a converted SELL handle (X, Y row-major, Y in original row
 * order): each output is a sequential fma chain over exactly the row's
 * stored entries in CRS order (slot_len bounds the chain, no padding term),
 * so Y equals the reference's row-major CRS SpMM
 * (spmm_cpu_crs_impl.hpp:50-89) bit for bit.  Chunks run long-first.
 */
int smax_hip_scs_apply_spmm(smax_hip_ctx ctx, int value_type, smax_hip_scs A,
                            int64_t k, const void *X, void *Y);

/* ------------------------------------------------ CRS -> SELL-C-sigma (GPU) */
typedef struct {
    int64_t C, sigma, n_rows, n_rows_padded, n_cols, n_chunks, n_elements,
        nnz, val_bytes;
    int32_t *chunk_ptr;     /* n_chunks + 1          */
    int32_t *chunk_lengths; /* n_chunks              */
    int32_t *col;           /* n_elements, padding 0 */
    void *val;              /* n_elements, padding 0 */
    int32_t *perm;          /* n_rows: perm[old] = new slot            */
    int32_t *slot_row;      /* n_rows_padded: slot -> old row (or pad) */
    int32_t *src_index;     /* n_elements: CRS nnz index, -1 = padding */
    int32_t *slot_len;      /* n_rows_padded: stored elements of each slot's
                               row (the sorted row lengths; 0 for padding) */
} smax_hip_scs_info;

/*
 * Converts a device CRS matrix to SELL-C-sigma on the GPU, bit-identical to
 * SMAX::Utils::convert_crs_to_scs (src/utils/convert_formats.hpp:16-132),
 * including the order std::sort (libstdc++) leaves tied rows in inside each
 * sigma window.  val_bytes in {2, 4, 8}: values are copied bit for bit.
 * The handle owns every array in its info.
 */
int smax_hip_scs_from_crs(smax_hip_ctx ctx, int64_t n_rows, int64_t n_cols,
                          int64_t nnz, const int32_t *row_ptr,
                          const int32_t *col, const void *val,
                          int64_t val_bytes, int64_t C, int64_t sigma,
                          smax_hip_scs *out);

/*
 * A handle over SELL-C-sigma arrays the caller already holds on the device
 * (the 13 values SpMVKernel::register_A takes for a SELL matrix,
 * src/kernels/spmv.hpp:37-50, with val_bytes in {2, 4, 8}): the arrays are
 * copied, so the caller's may change or go away afterwards.  It derives what
 * smax_hip_scs_apply needs from them -- each slot's row from perm, each
 * slot's length (its elements up to the last one that is not a +0.0 value at
 * column 0: such trailing entries add exactly the padding term 0 * x[0], so
 * the product is unchanged bit for bit), the execution schedule and the tail
 * waves -- and refuses arrays whose chunk_ptr does not match
 * C * chunk_lengths, whose column indices leave [0, n_cols) or whose perm is
 * not a permutation into [0, n_rows_padded).  Only smax_hip_scs_apply,
 * _get_info, _get_tails and _destroy take such a handle (no CRS positions:
 * no value refresh, SDDMM or solvers; no SpMM).  No reference counterpart:
 * it is how the SMAX::Interface HIP platform runs registered SELL matrices.
 */
int smax_hip_scs_from_sell(smax_hip_ctx ctx, int64_t C, int64_t sigma, int64_t n_rows,
                           int64_t n_rows_padded, int64_t n_cols, int64_t n_chunks,
                           int64_t n_elements, int64_t nnz, const int32_t *chunk_ptr,
                           const int32_t *chunk_lengths, const int32_t *col, const void *val,
                           int64_t val_bytes, const int32_t *perm, smax_hip_scs *out);
int smax_hip_scs_get_info(smax_hip_scs scs, smax_hip_scs_info *info);
int smax_hip_scs_destroy(smax_hip_ctx ctx, smax_hip_scs scs);

/*
 * SpMV with a matrix built by smax_hip_scs_from_crs: the same result, bit for
 * bit, as smax_hip_scs_spmv (unpermuted = 0: y in SELL order, n_rows_padded
 * entries) or smax_hip_scs_spmv_unpermuted (unpermuted = 1).  It uses the
 * handle's slot lengths so each lane stops at its row's last stored element
 * (padding is neither read nor multiplied; a non-finite x[0] still reaches
 * every padded row as the reference's 0*x[0] term would), and runs in the
 * handle's execution schedule: long chunks first, contiguous chunk ranges per
 * XCD when sigma*16 <= n_rows_padded, and a capped occupancy for layouts with
 * little padding.  No reference counterpart.
 */
int smax_hip_scs_apply(smax_hip_ctx ctx, int value_type, smax_hip_scs A,
                       const void *x, void *y, int unpermuted);

/*
 * How smax_hip_scs_apply runs A: the slots moved out of chunks that mix rows
 * of very different lengths into tail waves of their own (n_slots; 0 when
 * the layout has none), the number of those waves, and the entries they
 * store (copies of those slots' elements, with padding).  A layout with
 * 16 <= C <= 256 (C a multiple of 8) gets tail waves when its padding exceeds
 * 5x its stored entries; in such a chunk the slots longer than its (C/8)-th
 * longest slot move when that slot is under half the chunk's length.
 * Results are unchanged bit for bit.  No reference counterpart.
 */
int smax_hip_scs_get_tails(smax_hip_scs A, int64_t *n_slots, int64_t *n_waves,
                           int64_t *n_entries);

/* Re-gathers the SELL values from a CRS value array of the same pattern
 * (val_sell[e] = crs_val[src_index[e]], 0 for padding).  Asynchronous, except
 * when a solver left rounded value copies on the handle: those are dropped
 * (after a stream synchronisation) and rebuilt by the next solve. */
int smax_hip_scs_refresh_values(smax_hip_ctx ctx, smax_hip_scs scs,
                                const void *crs_val);

/* ------------------------------------------------------------------- BCRS */
/*
 * Block SpMV: per block row r and block row-offset ix < b_h,
 * y[r*h_pad + ix] = sum_j sum_{jx < b_w} val[j*h_pad*w_pad + off] *
 *                   x[col[j]*w_pad + jx],
 * off = jx*h_pad + ix (block column-major) or ix*w_pad + jx (row-major).
 * Replaces naive_bcrs_spmv_cuda_launcher
 * (src/kernels/spmv/spmv_cuda/spmv_cuda_bcrs_impl.cu:178-262) and restates
 * naive_bcrs_spmv (src/kernels/spmv/spmv_cpu/spmv_cpu_bcrs_impl.hpp:19-85).
 * Each output is a sequential fma chain in (j, jx) order.
 */
int smax_hip_bcrs_spmv(smax_hip_ctx ctx, int value_type,
                       int64_t n_block_rows, int64_t b_h, int64_t b_w,
                       int64_t h_pad, int64_t w_pad, const int32_t *col,
                       const int32_t *row_ptr, const void *val,
                       int block_col_major, const void *x, void *y);

/* Block SpMM with k right-hand sides, X and Y row-major
 * (X[(c*w_pad + jx)*k + v], Y[(r*h_pad + ix)*k + v]).  No reference
 * counterpart: the reference applies BCRS SpMV once per vector. */
int smax_hip_bcrs_spmm(smax_hip_ctx ctx, int value_type,
                       int64_t n_block_rows, int64_t b_h, int64_t b_w,
                       int64_t h_pad, int64_t w_pad, const int32_t *col,
                       const int32_t *row_ptr, const void *val,
                       int block_col_major, int64_t k, const void *X, void *Y);

typedef struct {
    int64_t n_block_rows, n_block_cols, n_blocks, b_h, b_w, h_pad, w_pad,
        val_bytes, block_col_major;
    int32_t *col;       /* n_blocks         */
    int32_t *row_ptr;   /* n_block_rows + 1 */
    void *val;          /* n_blocks * h_pad * w_pad, zero filled */
    int32_t *src_block; /* transposes: block of the source matrix each block
                           copies (NULL for a conversion from CRS) */
} smax_hip_bcrs_info;

/* CRS -> BCRS on the GPU, identical to SMAX::Utils::convert_crs_to_bcrs
 * (src/utils/convert_formats.hpp:134-247).  Rejects b_h/b_w that do not
 * divide the shape, padding smaller than the block, and rows whose columns
 * are not ascending (the reference throws for each); also rejects column
 * indices outside [0, n_cols), which the reference would place outside
 * their block.  A column listed twice in a row keeps its later value. */
int smax_hip_bcrs_from_crs(smax_hip_ctx ctx, int64_t n_rows, int64_t n_cols,
                           int64_t nnz, const int32_t *row_ptr,
                           const int32_t *col, const void *val,
                           int64_t val_bytes, int64_t b_h, int64_t b_w,
                           int64_t h_pad, int64_t w_pad, int block_col_major,
                           smax_hip_bcrs *out);
int smax_hip_bcrs_get_info(smax_hip_bcrs bcrs, smax_hip_bcrs_info *info);
int smax_hip_bcrs_destroy(smax_hip_ctx ctx, smax_hip_bcrs bcrs);

/* ---------------------------------------------------------- differentiation */
/*
 * BCRS layout of A^T built once from a BCRS A: the block pattern is
 * transposed with a stable sort by block column (deterministic), and a
 * column-major h x w block read row-major with the dimensions swapped is the
 * transposed block, so A^T's blocks are A's blocks (bytes unchanged,
 * reordered) with b_h/b_w, h_pad/w_pad and the block layout flag swapped.
 * Applying smax_hip_bcrs_spmm to it computes A^T G.  No reference
 * counterpart.
 */
int smax_hip_bcrs_transpose(smax_hip_ctx ctx, smax_hip_bcrs A,
                            smax_hip_bcrs *out);

/* Re-copies a transpose's blocks from the source matrix's current values
 * (val of the BCRS A it was built from). */
int smax_hip_bcrs_refresh_values(smax_hip_ctx ctx, smax_hip_bcrs T,
                                 const void *src_val);

/*
 * SDDMM on a BCRS layout: for every block j of block row r and in-block
 * position (ix < b_h, jx < b_w),
 *   dvals[j*h_pad*w_pad + off(ix,jx)] =
 *       sum_v G[(r*h_pad+ix)*k + v] * X[(col[j]*w_pad+jx)*k + v],
 * the gradient of sum(Y o G) w.r.t. the block values for Y = A X.  Padding
 * positions are set to zero.  Fixed reduction tree over v.
 */
int smax_hip_bcrs_sddmm(smax_hip_ctx ctx, int value_type, smax_hip_bcrs A,
                        int64_t k, const void *G, const void *X, void *dvals);

/*
 * SELL-C-sigma layout of A^T built once from A's device CRS pattern, with a
 * stable counting sort by column (deterministic, no atomics).  The result's
 * src_index refers to A's CRS nnz positions, so smax_hip_scs_refresh_values
 * with A's CRS values keeps A^T current.  No reference counterpart.
 */
int smax_hip_scs_transpose_from_crs(smax_hip_ctx ctx, int64_t n_rows,
                                    int64_t n_cols, int64_t nnz,
                                    const int32_t *row_ptr, const int32_t *col,
                                    const void *val, int64_t val_bytes,
                                    int64_t C, int64_t sigma,
                                    smax_hip_scs *out);

/*
 * SDDMM on A's SELL layout: dvals[src_index[e]] = sum_v G[r*k+v] *
 * X[col[e]*k+v] for every stored (non-padding) element e of row r, i.e. the
 * gradient of sum(Y o W) w.r.t. A's CRS values when G = W and Y = A X.
 * G is n_rows x k, X is n_cols x k, both row-major, original row order.
 * Accumulates in the value type's accumulation precision with a tree that
 * depends only on k: g = clamp(P/4, 1, 64) partial sums (P the smallest
 * power of two >= k), partial l an fma chain over v in [l*q, min(k, l*q+q)),
 * q = ceil(k/g), then a butterfly over the g partials.  G and X aligned to
 * 16 bytes take a faster kernel with identical results.  No reference
 * counterpart.
 */
int smax_hip_scs_sddmm(smax_hip_ctx ctx, int value_type, smax_hip_scs A,
                       int64_t k, const void *G, const void *X, void *dvals);

/* ----------------------------------------------------------- CG solvers */
/* Column codings the solvers may use for their slot-space matrix.  Every
 * coding gives each element the same column, so the iterates do not depend on
 * the choice; narrower ones move fewer index bytes per product. */
typedef enum {
    SMAX_HIP_COLUMNS_AUTO = 0, /* the narrowest that fits: one-byte indices
                                  into per-chunk tables of slot offsets (C =
                                  64, at most 32 distinct offsets per chunk),
                                  else 16-bit offsets from the slot, else
                                  32-bit columns                            */
    SMAX_HIP_COLUMNS_32 = 1,   /* 32-bit slot-space columns                  */
    SMAX_HIP_COLUMNS_16 = 2,   /* 16-bit offsets when they fit, else 32-bit  */
    SMAX_HIP_COLUMNS_8 = 3     /* one-byte table indices: as a request, the
                                  same as AUTO; reported when they ran      */
} smax_hip_column_coding;

typedef struct {
    double rel_tol;         /* stop when ||b - A x||_2 / ||b||_2 <= rel_tol */
    int64_t max_iters;      /* total (inner) iteration cap                   */
    int64_t check_every;    /* iterations enqueued between host checks; the
                               device stops itself on convergence (>= 1)     */
    double inner_rel_tol;   /* IR-CG: residual reduction of the low-precision
                               iteration between two FP64 refinement steps  */
    int64_t max_inner;      /* IR-CG: iteration cap per refinement step      */
    int32_t inner_value_type; /* IR-CG: SMAX_HIP_F32 or SMAX_HIP_BF16        */
    int32_t keep_diag_f32;  /* IR-CG, BF16 storage: apply the diagonal from an
                               FP32 side array instead of the BF16 values     */
    int32_t restart_inner;  /* IR-CG: 0 = keep the low-precision Krylov space
                               across refinement steps (reliable updates),
                               1 = restart CG on each correction equation
                               A d = r (classical iterative refinement)     */
    int32_t column_coding;  /* smax_hip_column_coding (0: narrowest)        */
} smax_hip_cg_params;

typedef struct {
    int64_t iterations;      /* FP64 CG iterations, or IR-CG inner total    */
    int64_t outer_iterations;/* IR-CG refinement steps (0 for plain CG)     */
    double rel_residual;     /* final ||b - A x||_2/||b||_2, FP64, on GPU   */
    double seconds;          /* wall time of the solve (setup excluded)     */
    double setup_seconds;    /* one-time build of the solver's matrix copies */
    int32_t column_coding;   /* the coding the products used: 1, 2 or 3
                                (SMAX_HIP_COLUMNS_32 / _16 / _8)             */
} smax_hip_cg_report;

/*
 * Plain FP64 conjugate gradients on a SELL-C-sigma SPD matrix (built by
 * smax_hip_scs_from_crs with 8-byte values).  b and x are device FP64
 * vectors of n_rows entries in original order; x holds the initial guess
 * and receives the solution.  Deterministic (fixed-order reductions).
 * No reference counterpart (the reference ships Jacobi/Gauss-Seidel demos,
 * examples/applications/).
 */
int smax_hip_cg_f64(smax_hip_ctx ctx, smax_hip_scs A, const double *b,
                    double *x, const smax_hip_cg_params *params,
                    smax_hip_cg_report *report);

/*
 * Mixed-precision iterative-refinement CG.  The residual r = b - A x, the
 * solution update and all refinement-loop norms and dot products are FP64 on
 * the FP64 matrix; the iteration between refinement steps runs CG on a copy
 * of A's values stored in FP32 or BF16 (FP32 vectors and accumulation).
 * Default (restart_inner = 0): the low-precision CG keeps its search
 * direction across refinement steps, and a refinement step (x += correction
 * in FP64, r = b - A x in FP64, low-precision residual replaced by it) runs
 * whenever the low-precision residual has dropped by inner_rel_tol since the
 * last one ("reliable updates").  A refinement step whose FP64 residual is
 * no smaller than the previous one restarts the low-precision CG from it; a
 * second such step in a row switches the rest of the solve to classical
 * refinement.  restart_inner = 1: classical refinement from the start, a
 * fresh CG solve of A d = r to inner_rel_tol per step.
 */
int smax_hip_ircg(smax_hip_ctx ctx, smax_hip_scs A, const double *b,
                  double *x, const smax_hip_cg_params *params,
                  smax_hip_cg_report *report);

/* ------------------------------------------------------------ measurement */
/* Device copy and triad (a = b + s*c) over n doubles: the stream bandwidth
 * reference for roofline fractions.  Enqueued on the current stream. */
int smax_hip_stream_copy(smax_hip_ctx ctx, const double *src, double *dst,
                         int64_t n);
int smax_hip_stream_triad(smax_hip_ctx ctx, double *a, const double *b,
                          const double *c, double s, int64_t n);
/* The copy with loads_per_lane (1, 2, 4 or 8) independent 16-byte loads in
 * flight per lane before its stores: a second method for the stream peak
 * (src and dst 16-byte aligned).  No reference counterpart. */
int smax_hip_stream_copy_wide(smax_hip_ctx ctx, const double *src, double *dst,
                              int64_t n, int loads_per_lane);
/* Random-gather reference for SpMVs whose x accesses are scattered: lane i
 * (i < n) reads per_lane (8, 16 or 32) entries of x[n] at pseudo-random
 * positions of a window of `window` entries centred on i (shifted to stay
 * inside [0, n)) and writes their sum to y[i].  No index array is read, so
 * n * per_lane / time is the gather ceiling for that window.  No reference
 * counterpart (a measurement helper, like the two above). */
int smax_hip_gather_rate(smax_hip_ctx ctx, const double *x, double *y, int64_t n,
                         int64_t window, int per_lane);

#ifdef __cplusplus
}
#endif

#endif /* SMAX_HIP_H */
