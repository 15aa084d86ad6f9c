// Path-integral simulated quantum annealing (mirrors sqa_annealer.hpp:18-94).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "qanneal/backend.hpp"
#include "qanneal/sqa_observer.hpp"
#include "qanneal/sqa_schedule.hpp"
#include "qanneal/sqa_state.hpp"
#include "qanneal/state.hpp"

namespace qanneal {

struct SQAResult {
    State best_state;
    double best_energy = 0.0;
    std::vector<double> energy_trace;
    std::vector<double> j_perp_trace;
    std::vector<double> chi_B_trace;
    double calibrated_eps_tilde = 0.0;
    double j_perp_start = 0.0;
    double resolved_j_perp_end = 0.0;
    double final_j_perp = 0.0;
};

class SQAAnnealer {
public:
    SQAAnnealer(const Hamiltonian &hamiltonian, SQASchedule schedule, std::size_t trotter_slices,
                std::size_t replicas = 1);
    SQAAnnealer(std::shared_ptr<Backend> backend, SQASchedule schedule, std::size_t trotter_slices,
                std::size_t replicas = 1);

    void set_seed(std::uint64_t seed);

    SQAResult run(std::size_t sweeps_per_beta, std::size_t worldline_sweeps, std::size_t cluster_sweeps = 0,
                  std::size_t continuous_time_slices = 0, SQAObserver *observer = nullptr);

    // Bond-susceptibility adaptive J_perp schedule (sqa_annealer.cpp:499-952) on the device.
    SQAResult run_optimal(double beta, double j_perp_start, double j_perp_end, double eps_tilde, double alpha,
                          std::size_t num_steps, std::size_t sweeps_per_step, std::size_t worldline_sweeps = 0,
                          std::size_t cluster_sweeps = 0, std::size_t calib_probes = 12,
                          std::size_t calib_sweeps = 10, const std::string &debug_csv_path = "",
                          double beta_ramp_fraction = 0.3, double beta_ramp_start = 0.1);

    // Last device run's diagnostics: sweep-kernel milliseconds and slice-move acceptance (FP32).
    double last_kernel_ms() const { return last_kernel_ms_; }
    double last_accept_rate() const { return last_accept_rate_; }

private:
    std::shared_ptr<Backend> backend_;
    SQASchedule schedule_;
    std::size_t slices_ = 0;
    std::size_t replicas_ = 0;
    std::mt19937_64 rng_;
    double last_kernel_ms_ = 0.0;
    double last_accept_rate_ = 0.0;
};

}  // namespace qanneal
