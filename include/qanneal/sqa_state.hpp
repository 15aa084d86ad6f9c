// Replica x Trotter-slice x spin lattice (mirrors sqa_state.hpp:13-86): one flat int8 buffer,
// index (replica * slices + slice) * spins + spin.
#pragma once

#include <cstddef>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <vector>

#include "qanneal/state.hpp"

namespace qanneal {

class SQAState {
public:
    SQAState() = default;
    SQAState(std::size_t replicas, std::size_t slices, std::size_t spins)
        : R_(replicas), P_(slices), N_(spins), buf_(replicas * slices * spins, 1) {
        if (!R_ || !P_ || !N_) throw std::invalid_argument("SQAState dimensions must be > 0.");
    }

    std::size_t replicas() const { return R_; }
    std::size_t slices() const { return P_; }
    std::size_t spins() const { return N_; }

    int8_t &at(std::size_t r, std::size_t t, std::size_t i) { return buf_[offset(r, t, i)]; }
    const int8_t &at(std::size_t r, std::size_t t, std::size_t i) const { return buf_[offset(r, t, i)]; }
    int8_t *slice_ptr(std::size_t r, std::size_t t) { return &buf_[offset(r, t, 0)]; }
    const int8_t *slice_ptr(std::size_t r, std::size_t t) const { return &buf_[offset(r, t, 0)]; }
    int8_t *data() { return buf_.data(); }
    const int8_t *data() const { return buf_.data(); }

    State slice_state(std::size_t r, std::size_t t) const {
        State s(N_);
        const int8_t *p = slice_ptr(r, t);
        for (std::size_t i = 0; i < N_; ++i) s[i] = p[i];
        return s;
    }
    void set_slice_state(std::size_t r, std::size_t t, const State &s) {
        if (s.size() != N_) throw std::invalid_argument("State size mismatch.");
        int8_t *p = slice_ptr(r, t);
        for (std::size_t i = 0; i < N_; ++i) p[i] = s[i];
    }

    // One uniform_int_distribution<int>(0,1) draw per entry in flat order (sqa_state.hpp:64-72).
    template <class URNG>
    static SQAState random(std::size_t replicas, std::size_t slices, std::size_t spins, URNG &rng) {
        SQAState st(replicas, slices, spins);
        std::uniform_int_distribution<int> coin(0, 1);
        for (auto &v : st.buf_) v = coin(rng) ? 1 : -1;
        return st;
    }

private:
    std::size_t R_ = 0, P_ = 0, N_ = 0;
    std::vector<int8_t> buf_;

    std::size_t offset(std::size_t r, std::size_t t, std::size_t i) const {
        if (r >= R_ || t >= P_ || i >= N_) throw std::invalid_argument("SQAState index out of range.");
        return (r * P_ + t) * N_ + i;
    }
};

}  // namespace qanneal
