// Observer hooks (mirrors observer.hpp:9-29). Device runs call them from the host between launches:
// a run with an observer goes step by step (a SweepObserver: sweep by sweep) and hands over the
// state after each, with the trajectory the run has without an observer.
#pragma once

#include <cstddef>

#include "qanneal/state.hpp"

namespace qanneal {

class Observer {
public:
    virtual ~Observer() = default;
    virtual void record(std::size_t step, double beta, double energy, const State &state) = 0;
};

class SweepObserver {
public:
    virtual ~SweepObserver() = default;
    virtual void record_sweep(std::size_t step, std::size_t sweep, double beta, double energy, const State &state) = 0;
};

}  // namespace qanneal
