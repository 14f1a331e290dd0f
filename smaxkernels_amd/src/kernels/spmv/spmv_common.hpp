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
};

} // namespace SMAX::KERNELS::SPMV
