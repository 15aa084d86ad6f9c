// SQA observer hooks (mirrors sqa_observer.hpp:9-43). An SQAObserver is called after every step with
// the whole lattice; the run keeps its trajectory. An SQASweepObserver selects the reference's
// sequential tracing path (sqa_annealer.cpp:317-453: replicas one after another, the annealer's own
// engine, no field cache), replayed in FP64 on the device whatever the backend precision.
#pragma once

#include <cstddef>
#include <vector>

#include "qanneal/sqa_state.hpp"

namespace qanneal {

enum class SQASweepPhase : int { Slice = 0, Worldline = 1, Cluster = 2 };

class SQAObserver {
public:
    virtual ~SQAObserver() = default;
    virtual void record(std::size_t step, double beta, double gamma, double avg_energy, const SQAState &state) = 0;
};

class SQASweepObserver {
public:
    virtual ~SQASweepObserver() = default;
    virtual void record_sweep(std::size_t step, std::size_t replica, std::size_t sweep, SQASweepPhase phase,
                              double beta, double gamma, double avg_energy,
                              const std::vector<double> &replica_energies, const SQAState &state) = 0;
};

}  // namespace qanneal
