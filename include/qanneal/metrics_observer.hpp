// The reference's ready-made observers (mirrors metrics_observer.hpp:11-193): per-step energy and
// magnetization traces, and strided state snapshots per sweep.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "qanneal/metrics.hpp"
#include "qanneal/observer.hpp"
#include "qanneal/sqa_observer.hpp"

namespace qanneal {

// SA: the running energy and the mean spin after every step.
class MetricsObserver : public Observer {
public:
    std::vector<double> energy_trace;
    std::vector<double> magnetization_trace;

    void record(std::size_t, double, double energy, const State &state) override {
        energy_trace.push_back(energy);
        magnetization_trace.push_back(magnetization(state.spins.data(), state.size()));
    }
    void clear() {
        energy_trace.clear();
        magnetization_trace.clear();
    }
};

// SQA: the mean slice energy and the mean spin over the whole lattice after every step. The spin
// sum is accumulated in double, one spin at a time in (replica, slice, spin) order, as the
// reference does (exact: every partial sum is an integer).
class SQAMetricsObserver : public SQAObserver {
public:
    std::vector<double> energy_trace;
    std::vector<double> magnetization_trace;

    void record(std::size_t, double, double, double avg_energy, const SQAState &state) override {
        energy_trace.push_back(avg_energy);
        const std::size_t total = state.replicas() * state.slices() * state.spins();
        double sum = 0.0;
        const int8_t *p = state.data();
        for (std::size_t k = 0; k < total; ++k) sum += static_cast<double>(p[k]);
        magnetization_trace.push_back(total > 0 ? sum / static_cast<double>(total) : 0.0);
    }
    void clear() {
        energy_trace.clear();
        magnetization_trace.clear();
    }
};

// SA: a snapshot of (step, sweep, beta, running energy, spins) after every sweep whose index within
// its step is a multiple of stride; per-step records are ignored.
class StateTraceObserver : public Observer, public SweepObserver {
public:
    std::size_t stride = 1;
    std::vector<std::size_t> step_trace;
    std::vector<std::size_t> sweep_trace;
    std::vector<double> beta_trace;
    std::vector<double> energy_trace;
    std::vector<std::vector<int8_t>> state_trace;

    void record(std::size_t, double, double, const State &) override {}
    void record_sweep(std::size_t step, std::size_t sweep, double beta, double energy, const State &state) override {
        if (stride > 1 && sweep % stride != 0) return;
        step_trace.push_back(step);
        sweep_trace.push_back(sweep);
        beta_trace.push_back(beta);
        energy_trace.push_back(energy);
        state_trace.push_back(state.spins);
    }
    void clear() {
        step_trace.clear();
        sweep_trace.clear();
        beta_trace.clear();
        energy_trace.clear();
        state_trace.clear();
    }
};

// SQA: a snapshot after every sweep (of any phase) whose index is a multiple of stride: step,
// replica, sweep, phase, beta, gamma, mean slice energy, per-replica mean slice energies and the
// flattened lattice. Selecting it runs the sequential tracing path (see sqa_observer.hpp).
class SQAStateTraceObserver : public SQAObserver, public SQASweepObserver {
public:
    std::size_t stride = 1;
    std::size_t replicas = 0;
    std::size_t slices = 0;
    std::size_t spins = 0;
    std::vector<std::size_t> step_trace;
    std::vector<std::size_t> replica_trace;
    std::vector<std::size_t> sweep_trace;
    std::vector<int> phase_trace;
    std::vector<double> beta_trace;
    std::vector<double> gamma_trace;
    std::vector<double> avg_energy_trace;
    std::vector<std::vector<double>> replica_energy_trace;
    std::vector<std::vector<int8_t>> state_trace;

    void record(std::size_t, double, double, double, const SQAState &) override {}
    void record_sweep(std::size_t step, std::size_t replica, std::size_t sweep, SQASweepPhase phase, double beta,
                      double gamma, double avg_energy, const std::vector<double> &replica_energies,
                      const SQAState &state) override {
        if (stride > 1 && sweep % stride != 0) return;
        if (replicas == 0) {
            replicas = state.replicas();
            slices = state.slices();
            spins = state.spins();
        }
        step_trace.push_back(step);
        replica_trace.push_back(replica);
        sweep_trace.push_back(sweep);
        phase_trace.push_back(static_cast<int>(phase));
        beta_trace.push_back(beta);
        gamma_trace.push_back(gamma);
        avg_energy_trace.push_back(avg_energy);
        replica_energy_trace.push_back(replica_energies);
        const int8_t *p = state.data();
        state_trace.emplace_back(p, p + replicas * slices * spins);
    }
    void clear() {
        replicas = slices = spins = 0;
        step_trace.clear();
        replica_trace.clear();
        sweep_trace.clear();
        phase_trace.clear();
        beta_trace.clear();
        gamma_trace.clear();
        avg_energy_trace.clear();
        replica_energy_trace.clear();
        state_trace.clear();
    }
};

}  // namespace qanneal
