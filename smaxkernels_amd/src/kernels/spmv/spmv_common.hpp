// Arguments and flags of an SpMV kernel (src/kernels/spmv/spmv_common.hpp:8-48).
#pragma once

#include "../../common.hpp"

namespace SMAX::KERNELS::SPMV {

struct Args {
    SparseMatrix A;  // registered host arrays (not owned)
    DenseMatrix x, y;
};

struct Flags {
    SpMVType kernel_type = SpMVType::naive_thread_per_row;
    bool is_mat_scs = false;
    bool is_mat_bcrs = false;
    bool is_block_column_major = false;
    bool scs_original_order = false;  // SELL y in original row order
};

// Entries of x the product reads and of y it writes, per format
// (spmv_cuda_core.cuh:34-110 sizes the device copies the same way): CRS
// n_cols / n_rows, SELL n_cols / n_chunks * C (n_rows in original row
// order), BCRS n_cols * w_pad / n_rows * h_pad.
inline void operand_extents(const Args &a, const Flags &f, ULL &x_need, ULL &y_need) {
    if (f.is_mat_scs) {
        x_need = a.A.scs->n_cols;
        y_need = f.scs_original_order ? a.A.scs->n_rows : a.A.scs->n_chunks * a.A.scs->C;
    } else if (f.is_mat_bcrs) {
        x_need = a.A.bcrs->n_cols * a.A.bcrs->width_pad;
        y_need = a.A.bcrs->n_rows * a.A.bcrs->height_pad;
    } else {
        x_need = a.A.crs->n_cols;
        y_need = a.A.crs->n_rows;
    }
}

// Whether the registered x and y hold what a product at these element
// offsets reads and writes.
inline bool operands_fit(const Args &a, const Flags &f, ULL x_off, ULL y_off) {
    ULL x_need = 0, y_need = 0;
    operand_extents(a, f, x_need, y_need);
    return x_off + x_need <= a.x.n_rows && y_off + y_need <= a.y.n_rows;
}

} // namespace SMAX::KERNELS::SPMV
