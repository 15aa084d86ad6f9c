// Worldline-magnetisation susceptibility schedule, method "sqa_chi" (mirrors sqa_chi_annealer.hpp).
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

struct SQAChiResult {
    State best_state;
    double best_energy = 0.0;
    std::vector<double> energy_trace;
    std::vector<double> beta_schedule;
    std::vector<double> gamma_schedule;
    std::vector<double> s_schedule;
    std::vector<double> scan_s;
    std::vector<double> scan_gamma;
    std::vector<double> scan_chi_B;
    double s_star = 0.0;
    double gamma_star = 0.0;
    double j_perp_star = 0.0;
    double chi_floor = 0.0;
    double driver_A0 = 0.0;
};

class SQAChiAnnealer {
public:
    SQAChiAnnealer(const Hamiltonian &hamiltonian, std::size_t trotter_slices, std::size_t replicas = 1);
    SQAChiAnnealer(std::shared_ptr<Backend> backend, std::size_t trotter_slices, std::size_t replicas = 1);

    void set_seed(std::uint64_t seed);

    SQAChiResult run_chi(double beta, double gamma_start, double gamma_end, std::size_t num_steps,
                         std::size_t sweeps_per_step, std::size_t scan_points = 16, std::size_t scan_sweeps = 30,
                         std::size_t scan_burn = 10, double chi_floor_fraction = 1e-6, double driver_A0 = 0.0,
                         double beta_ramp_fraction = 0.3, double beta_ramp_start = 0.1,
                         const std::string &debug_csv_path = "");

    // run_chi with the pilot scan replaced by caller-supplied chi_B values on the same s grid
    // (the surrogate path); the schedule construction and production anneal are unchanged.
    SQAChiResult run_chi_with_profile(const std::vector<double> &chi_B, double beta, double gamma_start,
                                      double gamma_end, std::size_t num_steps, std::size_t sweeps_per_step,
                                      double chi_floor_fraction = 1e-6, double driver_A0 = 0.0,
                                      double beta_ramp_fraction = 0.3, double beta_ramp_start = 0.1);

private:
    std::shared_ptr<Backend> backend_;
    std::size_t slices_ = 0;
    std::size_t replicas_ = 0;
    std::mt19937_64 rng_;
};

}  // namespace qanneal
