// HIP platform core of the SpMV kernel: the counterpart of
// src/kernels/spmv/spmv_cuda.hpp (type dispatch) and
// src/kernels/spmv/spmv_cuda/spmv_cuda_core.cuh:11-205 (initialize copies A,
// x and y to the device, apply launches, finalize copies y back).
//
// Differences from the CUDA platform, all deliberate: device buffers are
// owned and reused by the kernel (no leak per initialize); apply enqueues on
// the kernel's stream without a device-wide synchronise (finalize waits);
// failures return 1 with a warning instead of exiting the process.
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>

#include "../../platforms/hip/hip.hpp"
#include "spmv_common.hpp"

namespace SMAX::KERNELS::SPMV::HIP {

using PLATFORM::HIP::DeviceBuffer;
using PLATFORM::HIP::report;

inline int hip_index_type(IntType t) { return static_cast<int>(t); }

inline int hip_value_type(FloatType t) {
    switch (t) {
    case FloatType::FLOAT64: return SMAX_HIP_F64;
    case FloatType::FLOAT32: return SMAX_HIP_F32;
    default: return SMAX_HIP_BF16;
    }
}

// bytes of one stored value / one vector entry
inline size_t value_bytes(FloatType t) {
    return t == FloatType::FLOAT64 ? 8 : t == FloatType::FLOAT32 ? 4 : 2;
}
inline size_t vector_bytes(FloatType t) { return t == FloatType::FLOAT64 ? 8 : 4; }

struct DeviceState {
    PLATFORM::HIP::Stream stream;
    DeviceBuffer a0, a1, a2, a3, a4;  // format arrays (see initialize)
    DeviceBuffer x, y;
    smax_hip_crs_plan plan = nullptr;
    const void *plan_row_ptr = nullptr;  // host row_ptr the plan was built for
    ULL plan_rows = 0, plan_nnz = 0;
    bool use_plan = false;               // CRS: row blocks chosen at initialize
    // SELL: a handle over the uploaded arrays (slot lengths, execution
    // schedule, tail waves), rebuilt at every initialize
    smax_hip_scs scs = nullptr;
    ~DeviceState() {
        if (plan) smax_hip_crs_plan_destroy(stream.get(), plan);
        if (scs) smax_hip_scs_destroy(stream.get(), scs);
    }
};

inline int check_types(KernelContext *k) {
    int st = smax_hip_check_types(hip_index_type(k->int_type),
                                  hip_value_type(k->float_type));
    return report(nullptr, st, "type check");
}

inline int initialize(Timers *timers, KernelContext *k, Args *args, Flags *flags,
                      DeviceState &d) {
    ScopedPhase phase(timers, "initialize");
    if (check_types(k) || d.stream.open()) return 1;
    smax_hip_ctx ctx = d.stream.get();
    const size_t vb = value_bytes(k->float_type), xb = vector_bytes(k->float_type);
    ULL x_size, y_size;
    int rc = 0;
    if (flags->is_mat_scs) {
        // The arrays go up as registered and become a handle (copied), which
        // derives each slot's length, the execution schedule and the tail
        // waves from them: rebuilt here, so it always matches the values
        // registered now.  The reference's SELL kernel never reads perm; a
        // matrix registered without one runs with the identity.
        const SCSMatrix &A = *args->A.scs;
        if (d.scs) smax_hip_scs_destroy(ctx, d.scs);
        d.scs = nullptr;
        rc |= d.a0.upload(ctx, A.chunk_ptr, 4 * (A.n_chunks + 1));
        rc |= d.a1.upload(ctx, A.chunk_lengths, 4 * A.n_chunks);
        rc |= d.a2.upload(ctx, A.col, 4 * A.n_elements);
        rc |= d.a3.upload(ctx, A.val, vb * A.n_elements);
        if (A.perm) {
            rc |= d.a4.upload(ctx, A.perm, 4 * A.n_rows);
        } else {
            std::vector<int32_t> id(A.n_rows);
            for (ULL r = 0; r < A.n_rows; ++r) id[r] = static_cast<int32_t>(r);
            rc |= d.a4.upload(ctx, id.data(), 4 * A.n_rows);
            if (!rc) rc |= report(ctx, smax_hip_sync(ctx), "synchronize");
        }
        x_size = A.n_cols;
        y_size = flags->scs_original_order ? A.n_rows : A.n_rows_padded;
        if (!rc)
            rc |= report(ctx,
                         smax_hip_scs_from_sell(ctx, A.C, A.sigma, A.n_rows, A.n_rows_padded,
                                                A.n_cols, A.n_chunks, A.n_elements, A.nnz,
                                                d.a0.as<int32_t>(), d.a1.as<int32_t>(),
                                                d.a2.as<int32_t>(), d.a3.get(),
                                                static_cast<int64_t>(vb), d.a4.as<int32_t>(),
                                                &d.scs),
                         "SELL handle");
        // the handle holds its own copies of the big arrays
        d.a2.release();
        d.a3.release();
    } else if (flags->is_mat_bcrs) {
        const BCRSMatrix &A = *args->A.bcrs;
        rc |= d.a0.upload(ctx, A.row_ptr, 4 * (A.n_rows + 1));
        rc |= d.a1.upload(ctx, A.col, 4 * A.nnz);
        rc |= d.a2.upload(ctx, A.val, vb * A.nnz * A.height_pad * A.width_pad);
        x_size = A.n_cols * A.width_pad;
        y_size = A.n_rows * A.height_pad;
    } else {
        const CRSMatrix &A = *args->A.crs;
        rc |= d.a0.upload(ctx, A.row_ptr, 4 * (A.n_rows + 1));
        rc |= d.a1.upload(ctx, A.col, 4 * A.nnz);
        rc |= d.a2.upload(ctx, A.val, vb * A.nnz);
        x_size = A.n_cols;
        y_size = A.n_rows;
        // Both CRS schedules give every row one lane's fma chain (bit-identical
        // results); they differ in how the entries reach the lanes.  The warp
        // types select the row blocks; naive_thread_per_row keeps the direct
        // lane-per-row loads for rows of even length (C1, 512^2 Laplacian:
        // 3.4 us against 5.8) and takes the row blocks when the longest row
        // exceeds 4x the mean (C2, Lomax rows: 0.78 ms against 1.37).
        bool blocked = flags->kernel_type != SpMVType::naive_thread_per_row;
        if (!blocked && A.n_rows > 0) {
            const int32_t *rp = static_cast<const int32_t *>(A.row_ptr);
            int64_t longest = 0;
            for (ULL r = 0; r < A.n_rows; ++r)
                longest = std::max<int64_t>(longest, static_cast<int64_t>(rp[r + 1]) - rp[r]);
            blocked = static_cast<double>(longest) > 4.0 * static_cast<double>(A.nnz) /
                                                         static_cast<double>(A.n_rows);
        }
        d.use_plan = blocked;
        if (!rc && blocked &&
            (!d.plan || d.plan_row_ptr != A.row_ptr || d.plan_rows != A.n_rows ||
             d.plan_nnz != A.nnz)) {
            if (d.plan) smax_hip_crs_plan_destroy(ctx, d.plan);
            d.plan = nullptr;
            rc |= report(ctx, smax_hip_crs_plan_create(ctx, A.n_rows, A.nnz,
                                                       d.a0.as<int32_t>(), &d.plan),
                         "CRS row-block plan");
            d.plan_row_ptr = A.row_ptr;
            d.plan_rows = A.n_rows;
            d.plan_nnz = A.nnz;
        }
    }
    x_size = std::max(x_size, args->x.n_rows);
    y_size = std::max(y_size, args->y.n_rows);
    rc |= d.x.upload(ctx, args->x.val, xb * x_size);
    rc |= d.y.upload(ctx, args->y.val, xb * y_size);
    return rc ? 1 : 0;
}

inline int apply(Timers *timers, KernelContext *k, Args *args, Flags *flags,
                 DeviceState &d, ULL x_off, ULL y_off) {
    ScopedPhase phase(timers, "apply");
    smax_hip_ctx ctx = d.stream.get();
    if (!ctx) {
        ErrorHandler::warning("[HIP] apply before initialize");
        return 1;
    }
    const int vt = hip_value_type(k->float_type);
    const size_t xb = vector_bytes(k->float_type);
    const void *x = static_cast<const char *>(d.x.get()) + xb * x_off;
    void *y = static_cast<char *>(d.y.get()) + xb * y_off;
    int st;
    if (flags->is_mat_scs) {
        // y in SELL order (n_rows_padded entries), as the reference writes it,
        // or in original row order (n_rows entries)
        st = smax_hip_scs_apply(ctx, vt, d.scs, x, y, flags->scs_original_order ? 1 : 0);
    } else if (flags->is_mat_bcrs) {
        const BCRSMatrix &A = *args->A.bcrs;
        st = smax_hip_bcrs_spmv(ctx, vt, A.n_rows, A.b_height, A.b_width,
                                A.height_pad, A.width_pad, d.a1.as<int32_t>(),
                                d.a0.as<int32_t>(), d.a2.get(),
                                flags->is_block_column_major ? 1 : 0, x, y);
    } else {
        const CRSMatrix &A = *args->A.crs;
        st = smax_hip_crs_spmv(ctx, vt, A.n_rows, A.n_cols, A.nnz,
                               d.a0.as<int32_t>(), d.a1.as<int32_t>(), d.a2.get(),
                               x, y, d.use_plan ? d.plan : nullptr);
    }
    return report(ctx, st, "SpMV launch");
}

inline int finalize(Timers *timers, KernelContext *k, Args *args, Flags *flags,
                    DeviceState &d) {
    ScopedPhase phase(timers, "finalize");
    smax_hip_ctx ctx = d.stream.get();
    if (!ctx) {
        ErrorHandler::warning("[HIP] finalize before initialize");
        return 1;
    }
    ULL y_size;  // spmv_cuda_core.cuh:190-200
    if (flags->is_mat_scs)
        y_size = args->A.scs->n_rows;
    else if (flags->is_mat_bcrs)
        y_size = args->A.bcrs->n_rows * args->A.bcrs->height_pad;
    else
        y_size = args->A.crs->n_rows;
    int rc = report(ctx, smax_hip_sync(ctx), "synchronize");
    if (!rc) rc = d.y.download(args->y.val, vector_bytes(k->float_type) * y_size);
    return rc;
}

} // namespace SMAX::KERNELS::SPMV::HIP
