// SQA observer hooks (mirrors sqa_observer.hpp); not called by device runs (see observer.hpp).
#pragma once

#include <cstddef>
#include <vector>

#include "qanneal/sqa_state.hpp"

namespace qanneal {

enum class SQASweepPhase { Slice, Cluster, Worldline };

class SQAObserver {
public:
    virtual ~SQAObserver() = default;
    virtual void record(std::size_t step, double beta, double gamma, double avg_energy, const SQAState &state) = 0;
};

class SQASweepObserver : public SQAObserver {
public:
    virtual void record_sweep(std::size_t step, std::size_t replica, std::size_t sweep, SQASweepPhase phase,
                              double beta, double gamma, double avg_energy,
                              const std::vector<double> &replica_energies, const SQAState &state) = 0;
};

}  // namespace qanneal
