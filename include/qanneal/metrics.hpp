// Spin-vector observables (mirrors metrics.hpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>

namespace qanneal {

inline double magnetization(const int8_t *spins, std::size_t n) {
    if (n == 0) return 0.0;
    long long s = 0;
    for (std::size_t i = 0; i < n; ++i) s += spins[i];
    return static_cast<double>(s) / static_cast<double>(n);
}

inline double overlap(const int8_t *a, const int8_t *b, std::size_t n) {
    if (n == 0) return 0.0;
    long long s = 0;
    for (std::size_t i = 0; i < n; ++i) s += static_cast<long long>(a[i]) * b[i];
    return static_cast<double>(s) / static_cast<double>(n);
}

}  // namespace qanneal
