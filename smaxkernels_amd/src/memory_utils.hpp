// Type-erased kernel arguments, as in the reference (src/memory_utils.hpp:14-64):
// integers are widened to (unsigned) long long, pointers kept as void* / void**.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <variant>
#include <vector>

namespace SMAX {

using Variant = std::variant<long long, unsigned long long, void *, void **>;
using ULL = unsigned long long int;

template <typename T> inline T as(void *ptr) { return static_cast<T>(ptr); }
template <typename T> inline T as(const void *ptr) { return static_cast<T>(ptr); }

// Metadata argument as ULL; negative values and pointers are rejected
// (src/memory_utils.hpp:38-64 throws "metadata value < 0").
inline ULL get_ull(const Variant &v) {
    return std::visit(
        [](auto &&x) -> ULL {
            using T = std::decay_t<decltype(x)>;
            if constexpr (std::is_pointer_v<T>) {
                throw std::runtime_error("This should never be called");
            } else {
                if constexpr (std::is_signed_v<T>)
                    if (x < 0) throw std::runtime_error("metadata value < 0");
                return static_cast<ULL>(x);
            }
        },
        v);
}

} // namespace SMAX
