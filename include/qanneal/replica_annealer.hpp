// Independent simulated-annealing replicas (mirrors replica_annealer.hpp:15-53). On the HIP
// backend every replica runs as one read of a single device launch.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <random>
#include <vector>

#include "qanneal/backend.hpp"
#include "qanneal/metrics.hpp"
#include "qanneal/schedule.hpp"
#include "qanneal/state.hpp"

namespace qanneal {

struct ReplicaResult {
    State best_state;
    double best_energy = 0.0;
    std::vector<double> energy_trace;         // running energy at each schedule step end
    std::vector<double> magnetization_trace;  // mean spin at each schedule step end
};

struct MultiAnnealResult {
    std::vector<ReplicaResult> replicas;
    State global_best_state;
    double global_best_energy = 0.0;
    std::vector<double> average_energy_trace;
    std::vector<double> average_magnetization_trace;
};

class ReplicaAnnealer {
public:
    // Without an explicit backend the run goes to the HIP backend (FP32).
    ReplicaAnnealer(const Hamiltonian &hamiltonian, AnnealSchedule schedule, std::size_t replicas);
    ReplicaAnnealer(std::shared_ptr<Backend> backend, AnnealSchedule schedule, std::size_t replicas);

    void set_seed(std::uint64_t seed);

    MultiAnnealResult run(std::size_t sweeps_per_beta);

private:
    std::shared_ptr<Backend> backend_;
    AnnealSchedule schedule_;
    std::size_t replicas_ = 0;
    std::mt19937_64 rng_;
};

}  // namespace qanneal
