// Core types of the SMAX interface (src/common.hpp of the reference), with
// PlatformType::HIP and FloatType::BFLOAT16 added.
#pragma once

#include <memory>

#include "error_handler.hpp"
#include "memory_utils.hpp"
#include "stopwatch.hpp"

namespace SMAX {

enum class KernelType { SPMV, SPMM, SPGEMM, SPTRSV, SPTRSM };

// src/common.hpp:15-19.  On HIP the BCRS SpMV schedule is fixed (LDS-staged
// blocks) and CRS SpMV uses naive_thread_per_row as lane-per-row and the
// warp variants as the LDS row-block schedule; all are bit-identical.
enum class SpMVType : int {
    naive_thread_per_row = 0,
    naive_warp_group = 1,
    naive_warp_shuffle = 2
};

enum class PlatformType { CPU, CUDA, HIP };
enum class IntType { INT16, INT32, INT64, UINT16, UINT32, UINT64 };
// BFLOAT16: values stored as bfloat16 bit patterns (uint16_t), vectors and
// accumulation in FLOAT32.
enum class FloatType { FLOAT32, FLOAT64, BFLOAT16 };

struct KernelContext {
    KernelType kernel_type;
    PlatformType platform_type;
    IntType int_type;
    FloatType float_type;
    KernelContext(KernelType kt, PlatformType pt, IntType it, FloatType ft)
        : kernel_type(kt), platform_type(pt), int_type(it), float_type(ft) {}
};

struct CRSMatrix {
    ULL n_rows = 0, n_cols = 0, nnz = 0;
    void *col = nullptr, *row_ptr = nullptr, *val = nullptr;
};

struct SCSMatrix {
    ULL C = 0, sigma = 0, n_rows = 0, n_rows_padded = 0, n_cols = 0,
        n_chunks = 0, n_elements = 0, nnz = 0;
    void *chunk_ptr = nullptr, *chunk_lengths = nullptr, *col = nullptr,
         *val = nullptr, *perm = nullptr;
};

struct BCRSMatrix {
    ULL n_rows = 0, n_cols = 0, nnz = 0, b_height = 0, b_width = 0,
        height_pad = 0, width_pad = 0;
    void *col = nullptr, *row_ptr = nullptr, *val = nullptr;
};

struct SparseMatrix {
    std::unique_ptr<CRSMatrix> crs;
    std::unique_ptr<BCRSMatrix> bcrs;
    std::unique_ptr<SCSMatrix> scs;
};

// User-owned dense operand (registered pointer, never freed by SMAX).
struct DenseMatrix {
    ULL n_rows = 0, n_cols = 0;
    void *val = nullptr;
};

struct UtilitiesContainer {
    int *lvl_ptr = nullptr;
    int n_levels = 0;
};

} // namespace SMAX
