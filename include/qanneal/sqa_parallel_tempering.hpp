// SQA parallel tempering over a (beta, Gamma) ladder (mirrors sqa_parallel_tempering.hpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "qanneal/backend.hpp"
#include "qanneal/state.hpp"

namespace qanneal {

struct SQAParallelTemperingResult {
    std::vector<State> final_states;
    std::vector<double> final_energies;
    State best_state;
    double best_energy = 0.0;
    std::vector<double> average_energy_trace;
    std::vector<double> swap_acceptance_trace;
    std::vector<double> j_perp_trace;
    std::vector<double> chi_B_trace;
    double calibrated_eps_tilde = 0.0;
    double j_perp_start = 0.0;
    double resolved_j_perp_end = 0.0;
    double final_j_perp = 0.0;
};

class SQAParallelTemperingAnnealer {
public:
    SQAParallelTemperingAnnealer(const Hamiltonian &hamiltonian, std::vector<double> betas,
                                 std::vector<double> gammas, std::size_t trotter_slices);
    SQAParallelTemperingAnnealer(std::shared_ptr<Backend> backend, std::vector<double> betas,
                                 std::vector<double> gammas, std::size_t trotter_slices);

    void set_seed(std::uint64_t seed);

    SQAParallelTemperingResult run(std::size_t sweeps_per_step, std::size_t worldline_sweeps, std::size_t steps,
                                   std::size_t swap_interval = 1, std::size_t cluster_sweeps = 0,
                                   std::size_t continuous_time_slices = 0);

    // Independent ladders in one device launch (FP32); element l is ladder l's result. With
    // step_betas / step_gammas (steps rows of one (beta, Gamma) pair per rung) every PT step runs
    // its own row (an annealed ladder); empty keeps the constructor's ladder at every step.
    std::vector<SQAParallelTemperingResult> run_ladders(std::size_t ladders, std::size_t sweeps_per_step,
                                                        std::size_t worldline_sweeps, std::size_t steps,
                                                        std::size_t swap_interval = 1,
                                                        std::size_t cluster_sweeps = 0,
                                                        const std::vector<std::vector<double>> &step_betas = {},
                                                        const std::vector<std::vector<double>> &step_gammas = {});

    // Adaptive shared-J_perp schedule (sqa_parallel_tempering.cpp:348-882) on the device.
    SQAParallelTemperingResult run_optimal(std::size_t num_steps, std::size_t sweeps_per_step,
                                           std::size_t worldline_sweeps, double eps_tilde,
                                           double alpha = 15.0 / 14.0, double j_perp_end = 0.0,
                                           std::size_t cluster_sweeps = 0, std::size_t swap_interval = 1,
                                           std::size_t continuous_time_slices = 0, std::size_t calib_probes = 12,
                                           std::size_t calib_sweeps = 10, const std::string &debug_csv_path = "");

private:
    std::shared_ptr<Backend> backend_;
    std::vector<double> betas_;
    std::vector<double> gammas_;
    std::size_t slices_ = 0;
    std::mt19937_64 rng_;
};

}  // namespace qanneal
