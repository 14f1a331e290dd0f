// SMAX::Interface — kernel registry and entry point (src/interface.hpp:21-48).
#pragma once

#include <memory>
#include <string>
#include <unordered_map>

#include "common.hpp"
#include "kernel.hpp"
#include "kernels/spmm.hpp"
#include "kernels/spmv.hpp"
#include "utils.hpp"

namespace SMAX {

class Interface {
    std::unique_ptr<UtilitiesContainer> uc;
    std::unordered_map<std::string, std::unique_ptr<Kernel>> kernels;

  public:
    std::unique_ptr<Timers> timers;
    std::unique_ptr<Utils> utils;

    Interface();
    ~Interface();

    // Refuses platforms this build does not provide with ErrorHandler::fatal
    // (a runtime_error), as interface.cpp:61-67 does for CUDA; this build
    // provides the HIP platform.
    int register_kernel(const std::string &name, KernelType kernel_type,
                        PlatformType platform = PlatformType::CPU,
                        IntType int_type = IntType::INT32,
                        FloatType float_type = FloatType::FLOAT64);

    // Throws (ErrorHandler::kernel_dne) when no kernel has that name.
    Kernel *kernel(const std::string &kernel_name);

    int get_n_levels() const { return uc->n_levels; }
};

} // namespace SMAX
