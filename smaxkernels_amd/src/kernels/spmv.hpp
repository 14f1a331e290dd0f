// SpMV kernel front-end: registration (CRS 6 args, Sell-C-sigma 13, BCRS 10),
// flags and the platform switch of src/kernels/spmv.hpp:12-201, with the CPU
// and HIP platforms behind initialize / apply / finalize.
#pragma once

#include <stdexcept>

#include "../kernel.hpp"
#include "spmv/spmv_common.hpp"
#include "spmv/spmv_cpu.hpp"
#include "spmv/spmv_hip.hpp"

namespace SMAX::KERNELS {

class SpMVKernel : public Kernel {
  public:
    SPMV::Args args;
    SPMV::Flags flags;
    std::unique_ptr<SPMV::HIP::DeviceState> dev;

    explicit SpMVKernel(std::unique_ptr<KernelContext> ctx)
        : Kernel(std::move(ctx)), dev(std::make_unique<SPMV::HIP::DeviceState>()) {}

    // Flags must be set before register_A (spmv.hpp:26,52).
    int _register_A(const std::vector<Variant> &a) override {
        if (flags.is_mat_scs) {
            if (a.size() != 13)
                throw std::runtime_error("SpMVKernel register_A expects 13 Sell-C-sigma args");
            auto m = std::make_unique<SCSMatrix>();
            m->C = get_ull(a[0]);
            m->sigma = get_ull(a[1]);
            m->n_rows = get_ull(a[2]);
            m->n_rows_padded = get_ull(a[3]);
            m->n_cols = get_ull(a[4]);
            m->n_chunks = get_ull(a[5]);
            m->n_elements = get_ull(a[6]);
            m->nnz = get_ull(a[7]);
            m->chunk_ptr = std::get<void *>(a[8]);
            m->chunk_lengths = std::get<void *>(a[9]);
            m->col = std::get<void *>(a[10]);
            m->val = std::get<void *>(a[11]);
            m->perm = std::get<void *>(a[12]);
            args.A.scs = std::move(m);
        } else if (flags.is_mat_bcrs) {
            if (a.size() != 10)
                throw std::runtime_error("SpMVKernel register_A expects 10 BCRS args");
            auto m = std::make_unique<BCRSMatrix>();
            m->n_rows = get_ull(a[0]);
            m->n_cols = get_ull(a[1]);
            m->nnz = get_ull(a[2]);
            m->b_height = get_ull(a[3]);
            m->b_width = get_ull(a[4]);
            m->height_pad = get_ull(a[5]);
            m->width_pad = get_ull(a[6]);
            m->col = std::get<void *>(a[7]);
            m->row_ptr = std::get<void *>(a[8]);
            m->val = std::get<void *>(a[9]);
            args.A.bcrs = std::move(m);
        } else {
            if (a.size() != 6)
                throw std::runtime_error("SpMVKernel register_A expects 6 CRS args");
            auto m = std::make_unique<CRSMatrix>();
            m->n_rows = get_ull(a[0]);
            m->n_cols = get_ull(a[1]);
            m->nnz = get_ull(a[2]);
            m->col = std::get<void *>(a[3]);
            m->row_ptr = std::get<void *>(a[4]);
            m->val = std::get<void *>(a[5]);
            args.A.crs = std::move(m);
        }
        return 0;
    }

    int _register_B(const std::vector<Variant> &a) override {
        if (a.size() != 2) throw std::runtime_error("SpMVKernel register_B expects 2 args");
        args.x.n_rows = get_ull(a[0]);
        args.x.val = std::get<void *>(a[1]);
        return 0;
    }

    int _register_C(const std::vector<Variant> &a) override {
        if (a.size() != 2) throw std::runtime_error("SpMVKernel register_C expects 2 args");
        args.y.n_rows = get_ull(a[0]);
        args.y.val = std::get<void *>(a[1]);
        return 0;
    }

    int set_kernel_implementation(SpMVType t) override { flags.kernel_type = t; return 0; }
    int set_mat_scs(bool f) override { flags.is_mat_scs = f; return 0; }
    int set_mat_bcrs(bool f) override { flags.is_mat_bcrs = f; return 0; }
    int set_block_column_major(bool f) override { flags.is_block_column_major = f; return 0; }
    int set_scs_original_order(bool f) override { flags.scs_original_order = f; return 0; }

    int swap_operands() override {
        std::swap(args.x, args.y);
        dev->x.swap(dev->y);
        return 0;
    }

    // CPU: the registered host arrays are used in place (initialize and
    // finalize have nothing to move); HIP: copy in, launch, copy out.
    int initialize(ULL, ULL, ULL) override {
        if (!registered() || !SPMV::operands_fit(args, flags, 0, 0)) {
            ErrorHandler::warning("SpMV: x or y is smaller than the product reads or writes");
            return 1;
        }
        switch (k_ctx->platform_type) {
        case PlatformType::CPU: {
            ScopedPhase phase(timers.get(), "initialize");
            return 0;
        }
        case PlatformType::HIP:
            return SPMV::HIP::initialize(timers.get(), k_ctx.get(), &args, &flags, *dev);
        default:
            return unsupported_platform();
        }
    }
    int apply(ULL, ULL x_offset, ULL y_offset) override {
        if (!registered() || !SPMV::operands_fit(args, flags, x_offset, y_offset)) {
            ErrorHandler::warning("SpMV: x or y is smaller than the product reads or writes");
            return 1;
        }
        switch (k_ctx->platform_type) {
        case PlatformType::CPU: {
            ScopedPhase phase(timers.get(), "apply");
            if (CPU::spmv_apply(*k_ctx, args, flags, x_offset, y_offset) == 0) return 0;
            ErrorHandler::warning("[CPU] SpMV: no kernel for the registered types");
            return 1;
        }
        case PlatformType::HIP:
            return SPMV::HIP::apply(timers.get(), k_ctx.get(), &args, &flags, *dev,
                                    x_offset, y_offset);
        default:
            return unsupported_platform();
        }
    }
    int finalize(ULL, ULL, ULL) override {
        switch (k_ctx->platform_type) {
        case PlatformType::CPU: {
            ScopedPhase phase(timers.get(), "finalize");
            return 0;
        }
        case PlatformType::HIP:
            return SPMV::HIP::finalize(timers.get(), k_ctx.get(), &args, &flags, *dev);
        default:
            return unsupported_platform();
        }
    }

  private:
    // the matrix of the registered format, and both vectors, are registered
    bool registered() const {
        const bool a = flags.is_mat_scs ? bool(args.A.scs)
                                        : flags.is_mat_bcrs ? bool(args.A.bcrs) : bool(args.A.crs);
        return a && args.x.val && args.y.val;
    }
    static int unsupported_platform() {
        std::cerr << "Error: Platform not supported\n";
        return 1;
    }
};

} // namespace SMAX::KERNELS
