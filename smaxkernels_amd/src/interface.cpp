#include "interface.hpp"

#include <iomanip>
#include <iostream>

namespace SMAX {

Interface::Interface()
    : uc(std::make_unique<UtilitiesContainer>()), timers(std::make_unique<Timers>()) {
    ErrorHandler::initialize_log(".smax_log");
    utils = std::make_unique<Utils>(kernels, uc.get());
}

Interface::~Interface() {
    kernels.clear();
    utils.reset();
    ErrorHandler::close_log();
}

Kernel *Interface::kernel(const std::string &kernel_name) {
    auto it = kernels.find(kernel_name);
    if (it == kernels.end()) ErrorHandler::kernel_dne(kernel_name);
    return it->second.get();
}

int Interface::register_kernel(const std::string &name, KernelType kernel_type,
                               PlatformType platform, IntType int_type,
                               FloatType float_type) {
    if (platform == PlatformType::CUDA)
        ErrorHandler::fatal("Cannot register CUDA kernel. SMAX not built with CUDA enabled.");
    auto ctx = std::make_unique<KernelContext>(kernel_type, platform, int_type, float_type);
    switch (kernel_type) {
    case KernelType::SPMV:
        kernels[name] = std::make_unique<KERNELS::SpMVKernel>(std::move(ctx));
        return 0;
    case KernelType::SPMM:
        kernels[name] = std::make_unique<KERNELS::SpMMKernel>(std::move(ctx));
        return 0;
    default:
        std::cerr << "Error: Kernel not supported\n";
        return 1;
    }
}

void Utils::print_timers() {
    int width = 0;
    for (auto &[name, k] : kernels)
        width = std::max(width, static_cast<int>(name.size()) + 18);
    width += 2;
    std::cout << std::scientific << std::setprecision(3);
    auto line = [&](const std::string &label, long double t) {
        std::cout << std::left << std::setw(width) << label << std::right
                  << std::setw(12) << static_cast<double>(t) << " [s]\n";
    };
    for (auto &[name, k] : kernels) {
        const long double i = k->timers->get("initialize")->get_wtime();
        const long double a = k->timers->get("apply")->get_wtime();
        const long double f = k->timers->get("finalize")->get_wtime();
        std::cout << "+" << std::string(width + 14, '-') << "+\n";
        line("Total time in '" + name + "':", i + a + f);
        line("| Initialize time:", i);
        line("| Apply time:", a);
        line("| Finalize time:", f);
        std::cout << "+" << std::string(width + 14, '-') << "+\n";
    }
}

} // namespace SMAX
