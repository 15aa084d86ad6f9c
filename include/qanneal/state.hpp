// Spin state (mirrors the reference's state.hpp:10-50 interface).
#pragma once

#include <cstddef>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <vector>

namespace qanneal {

struct State {
    std::vector<int8_t> spins;  // each entry is -1 or +1

    State() = default;
    explicit State(std::size_t n) : spins(n, 1) {}

    std::size_t size() const { return spins.size(); }
    int8_t &operator[](std::size_t k) { return spins[k]; }
    const int8_t &operator[](std::size_t k) const { return spins[k]; }

    // One uniform_int_distribution<int>(0,1) draw per spin, in index order (state.hpp:22-30).
    template <class URNG>
    static State random(std::size_t n, URNG &rng) {
        State st(n);
        std::uniform_int_distribution<int> coin(0, 1);
        for (auto &v : st.spins) v = coin(rng) ? 1 : -1;
        return st;
    }
};

inline void validate_spins(const int8_t *spins, std::size_t n) {
    for (std::size_t k = 0; k < n; ++k)
        if (spins[k] != 1 && spins[k] != -1) throw std::invalid_argument("State spins must be -1 or +1.");
}

inline void validate_spins(const State &s) { validate_spins(s.spins.data(), s.size()); }

}  // namespace qanneal
