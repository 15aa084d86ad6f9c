// Simulated annealing (mirrors annealer.hpp:9-37). On the HIP backend run() executes on the GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <random>
#include <vector>

#include "qanneal/backend.hpp"
#include "qanneal/observer.hpp"
#include "qanneal/schedule.hpp"
#include "qanneal/state.hpp"

namespace qanneal {

struct AnnealResult {
    State best_state;
    double best_energy = 0.0;
    std::vector<double> energy_trace;
};

class Annealer {
public:
    // Without an explicit backend the run goes to the HIP backend (FP32).
    explicit Annealer(const Hamiltonian &hamiltonian, AnnealSchedule schedule);
    explicit Annealer(std::shared_ptr<Backend> backend, AnnealSchedule schedule);

    void set_seed(std::uint64_t seed);

    AnnealResult run(std::size_t sweeps_per_beta, Observer *observer = nullptr);

    // `reads` independent chains in one device launch. Validation backend: read r replays the
    // engine state the reference has after set_seed(seed + r) (what solve() does per read). FP32
    // backend: the launch's Philox stream is keyed by (seed, read index) -- reproducible, but not
    // the stream of a separate set_seed(seed + r) run.
    std::vector<AnnealResult> run_reads(std::size_t sweeps_per_beta, std::size_t reads, std::uint64_t seed);

private:
    std::shared_ptr<Backend> backend_;
    AnnealSchedule schedule_;
    std::mt19937_64 rng_;
};

}  // namespace qanneal
