// SpMM kernel front-end (src/kernels/spmm.hpp:11-106): CRS A times a block
// of k vectors.  The reference runs SpMM on the CPU only ("Platform not
// supported" otherwise, spmm.hpp:75-84); here the CPU and HIP platforms run
// it, with the same results bit for bit.
#pragma once

#include <algorithm>
#include <stdexcept>

#include "../kernel.hpp"
#include "spmv/spmv_cpu.hpp"
#include "spmv/spmv_hip.hpp"

namespace SMAX::KERNELS {

namespace SPMM {
struct Args {
    SparseMatrix A;
    DenseMatrix X, Y;
};
struct Flags {
    bool vec_row_major = false;  // spmm_common.hpp:34-36
};
} // namespace SPMM

class SpMMKernel : public Kernel {
  public:
    SPMM::Args args;
    SPMM::Flags flags;
    PLATFORM::HIP::Stream stream;
    PLATFORM::HIP::DeviceBuffer row_ptr, col, val, X, Y;
    // row execution order for row-major products, keyed by the host row_ptr
    smax_hip_crs_plan plan = nullptr;
    const void *plan_row_ptr = nullptr;
    ULL plan_rows = 0, plan_nnz = 0;

    explicit SpMMKernel(std::unique_ptr<KernelContext> ctx) : Kernel(std::move(ctx)) {}
    ~SpMMKernel() override {
        if (plan) smax_hip_crs_plan_destroy(stream.get(), plan);
    }

    int _register_A(const std::vector<Variant> &a) override {
        if (a.size() != 6) throw std::runtime_error("SpMMKernel register_A expects 6 args");
        auto m = std::make_unique<CRSMatrix>();
        m->n_rows = get_ull(a[0]);
        m->n_cols = get_ull(a[1]);
        m->nnz = get_ull(a[2]);
        m->col = std::get<void *>(a[3]);
        m->row_ptr = std::get<void *>(a[4]);
        m->val = std::get<void *>(a[5]);
        args.A.crs = std::move(m);
        return 0;
    }
    int _register_B(const std::vector<Variant> &a) override {
        if (a.size() != 3) throw std::runtime_error("SpMMKernel register_B expects 3 args");
        args.X.n_rows = get_ull(a[0]);
        args.X.n_cols = get_ull(a[1]);
        args.X.val = std::get<void *>(a[2]);
        return 0;
    }
    int _register_C(const std::vector<Variant> &a) override {
        if (a.size() != 3) throw std::runtime_error("SpMMKernel register_C expects 3 args");
        args.Y.n_rows = get_ull(a[0]);
        args.Y.n_cols = get_ull(a[1]);
        args.Y.val = std::get<void *>(a[2]);
        return 0;
    }
    int set_vec_row_major(bool f) override { flags.vec_row_major = f; return 0; }
    int swap_operands() override {
        std::swap(args.X, args.Y);
        X.swap(Y);
        return 0;
    }

    int initialize(ULL, ULL, ULL) override {
        if (!platform_ok()) return 1;
        ScopedPhase phase(timers.get(), "initialize");
        if (k_ctx->platform_type == PlatformType::CPU) return 0;  // host arrays in place
        if (SPMV::HIP::check_types(k_ctx.get()) || stream.open()) return 1;
        smax_hip_ctx c = stream.get();
        const CRSMatrix &A = *args.A.crs;
        const size_t vb = SPMV::HIP::value_bytes(k_ctx->float_type);
        const size_t xb = SPMV::HIP::vector_bytes(k_ctx->float_type);
        int rc = row_ptr.upload(c, A.row_ptr, 4 * (A.n_rows + 1));
        rc |= col.upload(c, A.col, 4 * A.nnz);
        rc |= val.upload(c, A.val, vb * A.nnz);
        rc |= X.upload(c, args.X.val, xb * args.X.n_rows * args.X.n_cols);
        rc |= Y.upload(c, args.Y.val, xb * args.Y.n_rows * args.Y.n_cols);
        if (!rc &&
            (!plan || plan_row_ptr != A.row_ptr || plan_rows != A.n_rows || plan_nnz != A.nnz)) {
            if (plan) smax_hip_crs_plan_destroy(c, plan);
            plan = nullptr;
            rc |= PLATFORM::HIP::report(
                c, smax_hip_crs_plan_create(c, A.n_rows, A.nnz, row_ptr.as<int32_t>(), &plan),
                "CRS SpMM plan");
            plan_row_ptr = A.row_ptr;
            plan_rows = A.n_rows;
            plan_nnz = A.nnz;
        }
        return rc ? 1 : 0;
    }

    int apply(ULL, ULL X_offset, ULL Y_offset) override {
        if (!platform_ok()) return 1;
        ScopedPhase phase(timers.get(), "apply");
        const CRSMatrix &A = *args.A.crs;
        const ULL k = args.X.n_cols;
        const size_t xb = SPMV::HIP::vector_bytes(k_ctx->float_type);
        // column-major operands use A's n_rows as leading dimension, as the
        // reference does (spmm_cpu_crs_impl.hpp:29,41)
        const long long ld = flags.vec_row_major ? static_cast<long long>(k)
                                                 : static_cast<long long>(A.n_rows);
        // the last element either operand's addressing reaches must lie in the
        // registered buffer (column-major with n_cols != n_rows can overrun)
        const ULL x_need = flags.vec_row_major ? A.n_cols * k : (k - 1) * ld + A.n_cols;
        const ULL y_need = flags.vec_row_major ? A.n_rows * k : (k - 1) * ld + A.n_rows;
        if (k == 0 || X_offset + x_need > args.X.n_rows * args.X.n_cols ||
            Y_offset + y_need > args.Y.n_rows * args.Y.n_cols) {
            ErrorHandler::warning("SpMM: X or Y is smaller than the product addresses");
            return 1;
        }
        if (k_ctx->platform_type == PlatformType::CPU) {
            KERNELS::CPU::SpMMApply f{A, static_cast<const char *>(args.X.val) + xb * X_offset,
                                      static_cast<char *>(args.Y.val) + xb * Y_offset, k,
                                      flags.vec_row_major};
            if (KERNELS::CPU::dispatch(*k_ctx, f) == 0) return 0;
            ErrorHandler::warning("[CPU] SpMM: no kernel for the registered types");
            return 1;
        }
        int st = smax_hip_crs_spmm(
            stream.get(), SPMV::HIP::hip_value_type(k_ctx->float_type), A.n_rows,
            A.n_cols, A.nnz, row_ptr.as<int32_t>(), col.as<int32_t>(), val.get(), k,
            flags.vec_row_major ? 1 : 0,
            static_cast<const char *>(X.get()) + xb * X_offset, ld,
            static_cast<char *>(Y.get()) + xb * Y_offset, ld,
            plan);
        return PLATFORM::HIP::report(stream.get(), st, "SpMM launch");
    }

    int finalize(ULL, ULL, ULL) override {
        if (!platform_ok()) return 1;
        ScopedPhase phase(timers.get(), "finalize");
        if (k_ctx->platform_type == PlatformType::CPU) return 0;
        int rc = PLATFORM::HIP::report(stream.get(), smax_hip_sync(stream.get()),
                                       "synchronize");
        if (!rc)
            rc = Y.download(args.Y.val, SPMV::HIP::vector_bytes(k_ctx->float_type) *
                                            args.Y.n_rows * args.Y.n_cols);
        return rc;
    }

  private:
    bool platform_ok() const {
        if (k_ctx->platform_type == PlatformType::HIP ||
            k_ctx->platform_type == PlatformType::CPU)
            return true;
        std::cerr << "Error: Platform not supported\n";
        return false;
    }
};

} // namespace SMAX::KERNELS
