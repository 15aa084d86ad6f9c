// TEST INFRASTRUCTURE ONLY. Implements oracle/orc_abi.h by calling the reference's own classes.
// oracle/Makefile compiles this file together with the reference sources (where they lie under
// /root/reference) into oracle/_ref/libqanneal_ref.so; it is the golden-vector source and the
// "reference" CPU baseline. No reference source is copied into this repository.
#include "orc_abi.h"

#include <cstring>
#include <stdexcept>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "qanneal/annealer.hpp"
#include "qanneal/parallel_tempering.hpp"
#include "qanneal/replica_annealer.hpp"
#include "qanneal/dense_ising.hpp"
#include "qanneal/higher_order_ising.hpp"
#include "qanneal/metrics.hpp"
#include "qanneal/observer.hpp"
#include "qanneal/sqa_observer.hpp"
#include "qanneal/qubo.hpp"
#include "qanneal/schedule.hpp"
#include "qanneal/sparse_ising.hpp"
#include "qanneal/sqa_annealer.hpp"
#include "qanneal/sqa_chi_annealer.hpp"
#include "qanneal/sqa_parallel_tempering.hpp"
#include "qanneal/sqa_schedule.hpp"

namespace {

thread_local std::string g_err;

std::unique_ptr<qanneal::Hamiltonian> build(const orc_problem *p) {
    const std::size_t n = static_cast<std::size_t>(p->n);
    if (p->kind == ORC_DENSE) {
        std::vector<double> h(p->h, p->h + n), J(p->J, p->J + n * n);
        return std::make_unique<qanneal::DenseIsing>(std::move(h), std::move(J), n, p->c);
    }
    if (p->kind == ORC_SPARSE) {
        std::vector<double> h(p->h, p->h + n);
        std::vector<qanneal::SparseEdge> edges;
        for (int64_t e = 0; e < p->nedges; ++e)
            edges.push_back({static_cast<std::size_t>(p->ei[e]), static_cast<std::size_t>(p->ej[e]), p->ev[e]});
        return std::make_unique<qanneal::SparseIsing>(std::move(h), std::move(edges), n, p->c);
    }
    std::vector<qanneal::HigherOrderIsing::RawTerm> terms;
    for (int64_t t = 0; t < p->nterms; ++t) {
        std::vector<std::size_t> vars;
        for (int64_t k = p->term_off[t]; k < p->term_off[t + 1]; ++k)
            vars.push_back(static_cast<std::size_t>(p->term_vars[k]));
        terms.emplace_back(std::move(vars), p->term_coeff[t]);
    }
    return std::make_unique<qanneal::HigherOrderIsing>(n, terms, p->c);
}

template <class F>
int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

void put_state(const qanneal::State &s, int8_t *out) {
    std::memcpy(out, s.spins.data(), s.spins.size());
}

}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_set_threads(int threads) {
#ifdef _OPENMP
    omp_set_num_threads(threads);
#endif
    (void)threads;
    return 0;
}

int orc_energy(const orc_problem *p, int64_t count, const int8_t *spins, double *out) {
    return guard([&] {
        auto ham = build(p);
        for (int64_t k = 0; k < count; ++k)
            out[k] = ham->energy(spins + k * p->n, static_cast<std::size_t>(p->n));
    });
}

int orc_delta_energy(const orc_problem *p, const int8_t *spins, int64_t flip, double *out) {
    return guard([&] {
        auto ham = build(p);
        *out = ham->delta_energy(spins, static_cast<std::size_t>(p->n), static_cast<std::size_t>(flip));
    });
}

int orc_local_fields(const orc_problem *p, const int8_t *spins, double *fields) {
    return guard([&] { build(p)->compute_local_fields(spins, static_cast<std::size_t>(p->n), fields); });
}

int orc_update_fields(const orc_problem *p, double *fields, const int8_t *new_spins, int64_t flip,
                      int8_t old_spin) {
    return guard([&] {
        build(p)->update_local_fields_after_flip(fields, new_spins, static_cast<std::size_t>(p->n),
                                                 static_cast<std::size_t>(flip), old_spin);
    });
}

int orc_qubo_to_ising(int64_t n, const double *Q, double *h, double *J, double *c) {
    return guard([&] {
        const std::size_t nn = static_cast<std::size_t>(n);
        qanneal::QUBO q(std::vector<double>(Q, Q + nn * nn), nn);
        qanneal::DenseIsing d = q.to_ising();
        std::memcpy(h, d.h().data(), nn * sizeof(double));
        std::memcpy(J, d.J().data(), nn * nn * sizeof(double));
        *c = d.constant();
    });
}

int orc_sa_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
               uint64_t seed, int8_t *best_state, double *best_energy, double *trace) {
    return guard([&] {
        auto ham = build(p);
        qanneal::AnnealSchedule sched;
        sched.betas.assign(betas, betas + steps);
        qanneal::Annealer ann(*ham, sched);
        ann.set_seed(seed);
        auto res = ann.run(static_cast<std::size_t>(sweeps_per_beta));
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        std::memcpy(trace, res.energy_trace.data(), res.energy_trace.size() * sizeof(double));
    });
}

namespace {
struct SaBoth : qanneal::Observer, qanneal::SweepObserver {
    double *mag, *se;
    int8_t *ss;
    std::size_t n, k = 0;
    void record(std::size_t step, double, double, const qanneal::State &s) override {
        mag[step] = qanneal::magnetization(s);
    }
    void record_sweep(std::size_t, std::size_t, double, double e, const qanneal::State &s) override {
        se[k] = e;
        std::memcpy(ss + k * n, s.spins.data(), n);
        ++k;
    }
};

struct SqaStep : qanneal::SQAObserver {
    double *mag;
    void record(std::size_t step, double, double, double, const qanneal::SQAState &st) override {
        double sum = 0.0;
        for (std::size_t r = 0; r < st.replicas(); ++r)
            for (std::size_t t = 0; t < st.slices(); ++t) {
                const int8_t *q = st.slice_ptr(r, t);
                for (std::size_t i = 0; i < st.spins(); ++i) sum += static_cast<double>(q[i]);
            }
        mag[step] = sum / static_cast<double>(st.replicas() * st.slices() * st.spins());
    }
};

struct SqaBoth : SqaStep, qanneal::SQASweepObserver {
    int64_t *meta;
    double *avg, *rep;
    int8_t *states;
    std::size_t k = 0;
    void record_sweep(std::size_t step, std::size_t replica, std::size_t sweep, qanneal::SQASweepPhase phase, double,
                      double, double a, const std::vector<double> &re, const qanneal::SQAState &st) override {
        meta[k * 4] = static_cast<int64_t>(step);
        meta[k * 4 + 1] = static_cast<int64_t>(replica);
        meta[k * 4 + 2] = static_cast<int64_t>(sweep);
        meta[k * 4 + 3] = static_cast<int64_t>(phase);
        avg[k] = a;
        std::memcpy(rep + k * re.size(), re.data(), re.size() * sizeof(double));
        const std::size_t cnt = st.replicas() * st.slices() * st.spins();
        std::memcpy(states + k * cnt, st.slice_ptr(0, 0), cnt);
        ++k;
    }
};
}  // namespace

int orc_sa_observe_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
                       uint64_t seed, int8_t *best_state, double *best_energy, double *trace, double *step_mag,
                       double *sweep_energies, int8_t *sweep_states) {
    return guard([&] {
        auto ham = build(p);
        qanneal::AnnealSchedule sched;
        sched.betas.assign(betas, betas + steps);
        qanneal::Annealer ann(*ham, sched);
        ann.set_seed(seed);
        SaBoth obs;
        obs.mag = step_mag;
        obs.se = sweep_energies;
        obs.ss = sweep_states;
        obs.n = static_cast<std::size_t>(p->n);
        auto res = ann.run(static_cast<std::size_t>(sweeps_per_beta), &obs);
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        std::memcpy(trace, res.energy_trace.data(), res.energy_trace.size() * sizeof(double));
    });
}

int orc_sqa_observe_run(const orc_problem *p, int64_t steps, const double *betas, const double *gammas,
                        int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                        int64_t worldline_sweeps, int64_t cluster_sweeps, int32_t tracing, uint64_t seed,
                        int8_t *best_state, double *best_energy, double *trace, double *step_mag,
                        int64_t *sweep_meta, double *sweep_avg, double *sweep_rep, int8_t *sweep_states) {
    return guard([&] {
        auto ham = build(p);
        auto sched = qanneal::SQASchedule::from_vectors(std::vector<double>(betas, betas + steps),
                                                        std::vector<double>(gammas, gammas + steps));
        qanneal::SQAAnnealer ann(*ham, sched, static_cast<std::size_t>(trotter_slices),
                                 static_cast<std::size_t>(replicas));
        ann.set_seed(seed);
        SqaBoth both;
        SqaStep plain;
        qanneal::SQAObserver *obs = &plain;
        plain.mag = step_mag;
        if (tracing) {
            both.mag = step_mag;
            both.meta = sweep_meta;
            both.avg = sweep_avg;
            both.rep = sweep_rep;
            both.states = sweep_states;
            obs = &both;
        }
        auto res = ann.run(static_cast<std::size_t>(sweeps_per_beta), static_cast<std::size_t>(worldline_sweeps),
                           static_cast<std::size_t>(cluster_sweeps), 0, obs);
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        std::memcpy(trace, res.energy_trace.data(), res.energy_trace.size() * sizeof(double));
    });
}

int orc_replica_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
                    int64_t replicas, uint64_t seed, int8_t *best_states, double *best_energies,
                    double *energy_traces, double *mag_traces, int8_t *global_best_state,
                    double *global_best_energy, double *avg_energy_trace, double *avg_mag_trace) {
    return guard([&] {
        auto ham = build(p);
        qanneal::AnnealSchedule sched;
        sched.betas.assign(betas, betas + steps);
        qanneal::ReplicaAnnealer ann(*ham, sched, static_cast<std::size_t>(replicas));
        ann.set_seed(seed);
        auto res = ann.run(static_cast<std::size_t>(sweeps_per_beta));
        for (int64_t r = 0; r < replicas; ++r) {
            const auto &rr = res.replicas[r];
            put_state(rr.best_state, best_states + r * p->n);
            best_energies[r] = rr.best_energy;
            std::memcpy(energy_traces + r * steps, rr.energy_trace.data(), steps * sizeof(double));
            std::memcpy(mag_traces + r * steps, rr.magnetization_trace.data(), steps * sizeof(double));
        }
        put_state(res.global_best_state, global_best_state);
        *global_best_energy = res.global_best_energy;
        std::memcpy(avg_energy_trace, res.average_energy_trace.data(), steps * sizeof(double));
        std::memcpy(avg_mag_trace, res.average_magnetization_trace.data(), steps * sizeof(double));
    });
}

int orc_pt_run(const orc_problem *p, int64_t replicas, const double *betas, int64_t sweeps_per_step,
               int64_t steps, int64_t swap_interval, uint64_t seed, int8_t *final_states,
               double *final_energies, int8_t *best_state, double *best_energy, double *avg_trace,
               double *swap_trace) {
    return guard([&] {
        auto ham = build(p);
        qanneal::ParallelTemperingAnnealer ann(*ham, std::vector<double>(betas, betas + replicas));
        ann.set_seed(seed);
        auto res = ann.run(static_cast<std::size_t>(sweeps_per_step), static_cast<std::size_t>(steps),
                           static_cast<std::size_t>(swap_interval));
        for (int64_t r = 0; r < replicas; ++r) {
            put_state(res.final_states[r], final_states + r * p->n);
            final_energies[r] = res.final_energies[r];
        }
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        std::memcpy(avg_trace, res.average_energy_trace.data(), steps * sizeof(double));
        std::memcpy(swap_trace, res.swap_acceptance_trace.data(), steps * sizeof(double));
    });
}

int orc_sqa_run(const orc_problem *p, int64_t steps, const double *betas, const double *gammas,
                int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                int64_t worldline_sweeps, int64_t cluster_sweeps, int64_t continuous_time_slices,
                uint64_t seed, int8_t *best_state, double *best_energy, double *trace) {
    return guard([&] {
        auto ham = build(p);
        auto sched = qanneal::SQASchedule::from_vectors(std::vector<double>(betas, betas + steps),
                                                        std::vector<double>(gammas, gammas + steps));
        qanneal::SQAAnnealer ann(*ham, sched, static_cast<std::size_t>(trotter_slices),
                                 static_cast<std::size_t>(replicas));
        ann.set_seed(seed);
        auto res = ann.run(static_cast<std::size_t>(sweeps_per_beta),
                           static_cast<std::size_t>(worldline_sweeps),
                           static_cast<std::size_t>(cluster_sweeps),
                           static_cast<std::size_t>(continuous_time_slices));
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        std::memcpy(trace, res.energy_trace.data(), res.energy_trace.size() * sizeof(double));
    });
}

int orc_sqa_run_repeat(const orc_problem *p, int64_t runs, int64_t steps, const double *betas,
                       const double *gammas, int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                       int64_t worldline_sweeps, int64_t cluster_sweeps, int64_t continuous_time_slices,
                       uint64_t seed, int8_t *best_state, double *best_energy, double *trace) {
    return guard([&] {
        if (runs <= 0) throw std::invalid_argument("runs must be > 0.");
        auto ham = build(p);
        auto sched = qanneal::SQASchedule::from_vectors(std::vector<double>(betas, betas + steps),
                                                        std::vector<double>(gammas, gammas + steps));
        qanneal::SQAAnnealer ann(*ham, sched, static_cast<std::size_t>(trotter_slices),
                                 static_cast<std::size_t>(replicas));
        ann.set_seed(seed);
        for (int64_t k = 0; k < runs; ++k) {
            auto res = ann.run(static_cast<std::size_t>(sweeps_per_beta), static_cast<std::size_t>(worldline_sweeps),
                               static_cast<std::size_t>(cluster_sweeps),
                               static_cast<std::size_t>(continuous_time_slices));
            put_state(res.best_state, best_state);
            *best_energy = res.best_energy;
            std::memcpy(trace, res.energy_trace.data(), res.energy_trace.size() * sizeof(double));
        }
    });
}

int orc_sqapt_run(const orc_problem *p, int64_t rungs, const double *betas, const double *gammas,
                  int64_t trotter_slices, int64_t sweeps_per_step, int64_t worldline_sweeps,
                  int64_t steps, int64_t swap_interval, int64_t cluster_sweeps,
                  int64_t continuous_time_slices, uint64_t seed, int8_t *final_states,
                  double *final_energies, int8_t *best_state, double *best_energy,
                  double *avg_trace, double *swap_trace) {
    return guard([&] {
        auto ham = build(p);
        qanneal::SQAParallelTemperingAnnealer ann(*ham, std::vector<double>(betas, betas + rungs),
                                                  std::vector<double>(gammas, gammas + rungs),
                                                  static_cast<std::size_t>(trotter_slices));
        ann.set_seed(seed);
        auto res = ann.run(static_cast<std::size_t>(sweeps_per_step),
                           static_cast<std::size_t>(worldline_sweeps), static_cast<std::size_t>(steps),
                           static_cast<std::size_t>(swap_interval),
                           static_cast<std::size_t>(cluster_sweeps),
                           static_cast<std::size_t>(continuous_time_slices));
        for (int64_t r = 0; r < rungs; ++r) {
            put_state(res.final_states[r], final_states + r * p->n);
            final_energies[r] = res.final_energies[r];
        }
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        std::memcpy(avg_trace, res.average_energy_trace.data(), steps * sizeof(double));
        std::memcpy(swap_trace, res.swap_acceptance_trace.data(), steps * sizeof(double));
    });
}

int orc_sqa_opt_run(const orc_problem *p, int64_t trotter_slices, int64_t replicas, double beta,
                    double j_perp_start, double j_perp_end, double eps_tilde, double alpha,
                    int64_t num_steps, int64_t sweeps_per_step, int64_t worldline_sweeps,
                    int64_t cluster_sweeps, int64_t calib_probes, int64_t calib_sweeps,
                    double beta_ramp_fraction, double beta_ramp_start, uint64_t seed,
                    int8_t *best_state, double *best_energy, double *energy_trace,
                    double *j_perp_trace, double *chi_trace, double *scalars) {
    return guard([&] {
        auto ham = build(p);
        auto sched = qanneal::SQASchedule::from_vectors({beta}, {1.0});  // unused by run_optimal
        qanneal::SQAAnnealer ann(*ham, sched, static_cast<std::size_t>(trotter_slices),
                                 static_cast<std::size_t>(replicas));
        ann.set_seed(seed);
        auto res = ann.run_optimal(beta, j_perp_start, j_perp_end, eps_tilde, alpha,
                                   static_cast<std::size_t>(num_steps), static_cast<std::size_t>(sweeps_per_step),
                                   static_cast<std::size_t>(worldline_sweeps),
                                   static_cast<std::size_t>(cluster_sweeps), static_cast<std::size_t>(calib_probes),
                                   static_cast<std::size_t>(calib_sweeps), "", beta_ramp_fraction, beta_ramp_start);
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        std::memcpy(energy_trace, res.energy_trace.data(), num_steps * sizeof(double));
        std::memcpy(j_perp_trace, res.j_perp_trace.data(), num_steps * sizeof(double));
        std::memcpy(chi_trace, res.chi_B_trace.data(), num_steps * sizeof(double));
        scalars[0] = res.calibrated_eps_tilde;
        scalars[1] = res.j_perp_start;
        scalars[2] = res.resolved_j_perp_end;
        scalars[3] = res.final_j_perp;
    });
}

int orc_sqapt_opt_run(const orc_problem *p, int64_t rungs, const double *betas, const double *gammas,
                      int64_t trotter_slices, int64_t num_steps, int64_t sweeps_per_step,
                      int64_t worldline_sweeps, double eps_tilde, double alpha, double j_perp_end,
                      int64_t cluster_sweeps, int64_t swap_interval, int64_t continuous_time_slices,
                      int64_t calib_probes, int64_t calib_sweeps, uint64_t seed, int8_t *final_states,
                      double *final_energies, int8_t *best_state, double *best_energy,
                      double *avg_trace, double *swap_trace, double *j_perp_trace, double *chi_trace,
                      double *scalars) {
    return guard([&] {
        auto ham = build(p);
        qanneal::SQAParallelTemperingAnnealer ann(*ham, std::vector<double>(betas, betas + rungs),
                                                  std::vector<double>(gammas, gammas + rungs),
                                                  static_cast<std::size_t>(trotter_slices));
        ann.set_seed(seed);
        auto res = ann.run_optimal(static_cast<std::size_t>(num_steps), static_cast<std::size_t>(sweeps_per_step),
                                   static_cast<std::size_t>(worldline_sweeps), eps_tilde, alpha, j_perp_end,
                                   static_cast<std::size_t>(cluster_sweeps), static_cast<std::size_t>(swap_interval),
                                   static_cast<std::size_t>(continuous_time_slices),
                                   static_cast<std::size_t>(calib_probes), static_cast<std::size_t>(calib_sweeps), "");
        for (int64_t r = 0; r < rungs; ++r) {
            put_state(res.final_states[r], final_states + r * p->n);
            final_energies[r] = res.final_energies[r];
        }
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        std::memcpy(avg_trace, res.average_energy_trace.data(), num_steps * sizeof(double));
        std::memcpy(swap_trace, res.swap_acceptance_trace.data(), num_steps * sizeof(double));
        std::memcpy(j_perp_trace, res.j_perp_trace.data(), num_steps * sizeof(double));
        std::memcpy(chi_trace, res.chi_B_trace.data(), num_steps * sizeof(double));
        scalars[0] = res.calibrated_eps_tilde;
        scalars[1] = res.j_perp_start;
        scalars[2] = res.resolved_j_perp_end;
        scalars[3] = res.final_j_perp;
    });
}

int orc_sqa_chi_run(const orc_problem *p, int64_t trotter_slices, int64_t replicas, double beta,
                    double gamma_start, double gamma_end, int64_t num_steps,
                    int64_t sweeps_per_step, int64_t scan_points, int64_t scan_sweeps,
                    int64_t scan_burn, double chi_floor_fraction, double driver_A0,
                    double beta_ramp_fraction, double beta_ramp_start, uint64_t seed,
                    int8_t *best_state, double *best_energy, double *trace, double *beta_sched,
                    double *gamma_sched, double *s_sched, double *scan_s, double *scan_gamma,
                    double *scan_chi, double *scalars) {
    return guard([&] {
        auto ham = build(p);
        qanneal::SQAChiAnnealer ann(*ham, static_cast<std::size_t>(trotter_slices),
                                    static_cast<std::size_t>(replicas));
        ann.set_seed(seed);
        auto res = ann.run_chi(beta, gamma_start, gamma_end, static_cast<std::size_t>(num_steps),
                               static_cast<std::size_t>(sweeps_per_step),
                               static_cast<std::size_t>(scan_points),
                               static_cast<std::size_t>(scan_sweeps),
                               static_cast<std::size_t>(scan_burn), chi_floor_fraction, driver_A0,
                               beta_ramp_fraction, beta_ramp_start);
        put_state(res.best_state, best_state);
        *best_energy = res.best_energy;
        const std::size_t ns = static_cast<std::size_t>(num_steps);
        const std::size_t sp = static_cast<std::size_t>(scan_points);
        std::memcpy(trace, res.energy_trace.data(), ns * sizeof(double));
        std::memcpy(beta_sched, res.beta_schedule.data(), ns * sizeof(double));
        std::memcpy(gamma_sched, res.gamma_schedule.data(), ns * sizeof(double));
        std::memcpy(s_sched, res.s_schedule.data(), ns * sizeof(double));
        std::memcpy(scan_s, res.scan_s.data(), sp * sizeof(double));
        std::memcpy(scan_gamma, res.scan_gamma.data(), sp * sizeof(double));
        std::memcpy(scan_chi, res.scan_chi_B.data(), sp * sizeof(double));
        scalars[0] = res.s_star;
        scalars[1] = res.gamma_star;
        scalars[2] = res.j_perp_star;
        scalars[3] = res.chi_floor;
        scalars[4] = res.driver_A0;
    });
}

}  // extern "C"
