// Kernel base class: argument packing, flag setters and run(), with the
// reference's interface and return conventions (src/kernel.hpp:7-127).
#pragma once

#include <iostream>
#include <type_traits>
#include <utility>
#include <vector>

#include "common.hpp"

#define SMAX_CHECK_ERROR(func, label)                                          \
    if (func) {                                                                \
        std::cerr << "Error in " << label << "\n";                             \
        return 1;                                                              \
    }

namespace SMAX {

class Kernel {
    template <typename T> static Variant wrap_arg(T **ptr) {
        return reinterpret_cast<void **>(ptr);
    }
    template <typename T> static Variant wrap_arg(T *ptr) {
        return static_cast<void *>(const_cast<std::remove_const_t<T> *>(ptr));
    }
    template <typename T>
    static std::enable_if_t<std::is_integral_v<T> && std::is_signed_v<T>, Variant>
    wrap_arg(T v) {
        return static_cast<long long>(v);
    }
    template <typename T>
    static std::enable_if_t<std::is_integral_v<T> && std::is_unsigned_v<T>, Variant>
    wrap_arg(T v) {
        return static_cast<unsigned long long>(v);
    }

  public:
    std::unique_ptr<KernelContext> k_ctx;
    std::unique_ptr<Timers> timers;

    explicit Kernel(std::unique_ptr<KernelContext> ctx)
        : k_ctx(std::move(ctx)), timers(std::make_unique<Timers>()) {
        timers->add("initialize");
        timers->add("apply");
        timers->add("finalize");
    }
    virtual ~Kernel() = default;

    virtual int initialize(ULL A_offset = 0, ULL B_offset = 0, ULL C_offset = 0) = 0;
    virtual int apply(ULL A_offset = 0, ULL B_offset = 0, ULL C_offset = 0) = 0;
    virtual int finalize(ULL A_offset = 0, ULL B_offset = 0, ULL C_offset = 0) = 0;
    virtual int _register_A(const std::vector<Variant> &args) = 0;
    virtual int _register_B(const std::vector<Variant> &args) = 0;
    virtual int _register_C(const std::vector<Variant> &args) = 0;

    // Flag setters return 1 when a kernel does not support them
    // (src/kernel.hpp:58-90).
    virtual int set_mat_perm(bool) { return unsupported("set_mat_perm"); }
    virtual int set_mat_upper_triang(bool) { return unsupported("set_mat_upper_triang"); }
    virtual int set_mat_lower_triang(bool) { return unsupported("set_mat_lower_triang"); }
    virtual int set_vec_row_major(bool) { return unsupported("set_vec_row_major"); }
    virtual int set_mat_scs(bool) { return unsupported("set_mat_scs"); }
    virtual int set_mat_bcrs(bool) { return unsupported("set_mat_bcrs"); }
    virtual int set_block_column_major(bool) { return unsupported("set_block_column_major"); }
    // SELL SpMV writes y in original row order (n_rows entries) instead of the
    // SELL order of the reference (n_rows_padded entries).  No reference
    // counterpart: the reference leaves y permuted for the caller to undo.
    virtual int set_scs_original_order(bool) { return unsupported("set_scs_original_order"); }
    virtual int set_kernel_implementation(SpMVType) {
        return unsupported("set_kernel_implementation(SpMVType)");
    }
    virtual int swap_operands() { return unsupported("swap_operands"); }

    int run(ULL A_offset = 0, ULL B_offset = 0, ULL C_offset = 0) {
        SMAX_CHECK_ERROR(initialize(A_offset, B_offset, C_offset), "initialize");
        SMAX_CHECK_ERROR(apply(A_offset, B_offset, C_offset), "apply");
        SMAX_CHECK_ERROR(finalize(A_offset, B_offset, C_offset), "finalize");
        return 0;
    }

    template <typename... Args> int register_A(Args &&...args) {
        return _register_A({wrap_arg(std::forward<Args>(args))...});
    }
    template <typename... Args> int register_B(Args &&...args) {
        return _register_B({wrap_arg(std::forward<Args>(args))...});
    }
    template <typename... Args> int register_C(Args &&...args) {
        return _register_C({wrap_arg(std::forward<Args>(args))...});
    }

  private:
    static int unsupported(const char *what) {
        std::cerr << what << " not supported for this kernel.\n";
        return 1;
    }
};

} // namespace SMAX
