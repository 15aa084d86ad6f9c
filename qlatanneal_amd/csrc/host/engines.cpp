// Annealer classes: argument checks as in the reference, then one whole run on the device through
// the C-ABI (include/qanneal_hip.h). For FP64 validation runs the host draws exactly what the
// reference's host code draws from the annealer's std::mt19937_64 (initial states, per-replica
// seeds, uniform streams) and advances the engine by what the device consumed, so consecutive
// runs and set_seed behave as in the reference.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <random>
#include <sstream>
#include <string>
#include <stdexcept>

#include "qanneal/annealer.hpp"
#include "qanneal/metrics_observer.hpp"
#include "qanneal/parallel_tempering.hpp"
#include "qanneal/replica_annealer.hpp"
#include "qanneal/sqa_annealer.hpp"
#include "qanneal/sqa_chi_annealer.hpp"
#include "qanneal/sqa_parallel_tempering.hpp"
#include "qanneal_hip.h"

namespace qanneal {

namespace {

HIPBackend &device_backend(const std::shared_ptr<Backend> &b) {
    if (!b->is_device())
        throw std::runtime_error(
            "this build runs the annealing sweep on the GPU only: construct the annealer with "
            "backend='hip' (or 'hip_fp64' for the FP64 validation replay)");
    return static_cast<HIPBackend &>(*b);
}

int mode_of(const HIPBackend &b) {
    return b.precision() == DevicePrecision::FP32 ? QA_MODE_FP32 : QA_MODE_FP64_VALIDATION;
}

std::vector<double> uniform_stream(const std::mt19937_64 &eng, std::size_t count) {
    std::mt19937_64 copy = eng;
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> out(count);
    for (auto &v : out) v = u(copy);
    return out;
}

State state_from(const int8_t *p, std::size_t n) {
    State s(n);
    std::memcpy(s.spins.data(), p, n);
    return s;
}

// Annealer::run with an observer (annealer.cpp:33-75): one launch per step (per sweep for a
// SweepObserver), each continuing from the previous launch's state and running energy, so the
// host can hand the state to the observer in between. In validation mode the chained launches
// consume the annealer's engine exactly as the single launch does; a launch's best replaces the
// run's best only when strictly lower, which is the reference's rule across the boundary too.
AnnealResult sa_observed(HIPBackend &hb, const AnnealSchedule &sched, std::mt19937_64 &rng, std::size_t spb,
                         Observer *observer) {
    auto *sweep_obs = dynamic_cast<SweepObserver *>(observer);
    const std::size_t n = hb.size(), steps = sched.size();
    const int mode = mode_of(hb);
    State state = State::random(n, rng);              // annealer.cpp:40
    double energy = hb.energy(state.spins.data(), n);  // :41
    const uint64_t key = mode == QA_MODE_FP32 ? rng() : 0;
    AnnealResult res;
    res.best_state = state;
    res.best_energy = energy;
    res.energy_trace.reserve(steps);
    std::vector<int8_t> bs(n), fs(n);
    double be = 0.0, tr = 0.0, fe = 0.0;
    int64_t used = 0;
    uint64_t launch = 0;
    const auto lease = hb.lease();
    for (std::size_t step = 0; step < steps; ++step) {
        const double beta = sched.betas[step];
        const std::size_t per = sweep_obs ? 1 : spb;
        for (std::size_t sw = 0; sw < spb; sw += per) {
            qa_sa_params prm{};
            prm.steps = 1;
            prm.betas = &sched.betas[step];
            prm.sweeps_per_beta = static_cast<int64_t>(per);
            prm.reads = 1;
            prm.mode = mode;
            prm.init_energies = &energy;
            std::vector<double> uni;
            if (mode == QA_MODE_FP32) {
                prm.seed = key;
                prm.init_states = state.spins.data();
                prm.counter_base = launch++;
            } else {
                uni = uniform_stream(rng, per * n);
                prm.init_spins = state.spins.data();
                prm.uniforms = uni.data();
                prm.uniforms_per_read = static_cast<int64_t>(uni.size());
            }
            qa_sa_result out{};
            out.best_states = bs.data();
            out.best_energies = &be;
            out.energy_trace = &tr;
            out.uniforms_used = &used;
            out.final_states = fs.data();
            out.final_energies = &fe;
            throw_on_status(qa_sa_run(lease.ctx, lease.problem, &prm, &out));
            if (mode != QA_MODE_FP32) rng.discard(static_cast<unsigned long long>(used));
            if (be < res.best_energy) {
                res.best_energy = be;
                res.best_state = state_from(bs.data(), n);
            }
            std::memcpy(state.spins.data(), fs.data(), n);
            energy = fe;
            if (sweep_obs) sweep_obs->record_sweep(step, sw, beta, energy, state);
        }
        res.energy_trace.push_back(energy);
        observer->record(step, beta, energy, state);
    }
    return res;
}

// SQAAnnealer::run's observer calls through the C-ABI callbacks.
struct SqaObserverCtx {
    SQAObserver *obs;
    SQASweepObserver *sweep;
    const SQASchedule *sched;
    SQAState state;
    std::vector<double> rep;
};

void sqa_on_step(void *user, int64_t step, double avg, const int8_t *spins) {
    auto &c = *static_cast<SqaObserverCtx *>(user);
    std::memcpy(c.state.data(), spins, c.state.replicas() * c.state.slices() * c.state.spins());
    c.obs->record(static_cast<std::size_t>(step), c.sched->betas[step], c.sched->gammas[step], avg, c.state);
}

void sqa_on_sweep(void *user, int64_t step, int64_t replica, int64_t sweep, int32_t phase, double avg,
                  const double *rep_e, const int8_t *spins) {
    auto &c = *static_cast<SqaObserverCtx *>(user);
    std::memcpy(c.state.data(), spins, c.state.replicas() * c.state.slices() * c.state.spins());
    c.rep.assign(rep_e, rep_e + c.state.replicas());
    c.sweep->record_sweep(static_cast<std::size_t>(step), static_cast<std::size_t>(replica),
                          static_cast<std::size_t>(sweep), static_cast<SQASweepPhase>(phase), c.sched->betas[step],
                          c.sched->gammas[step], avg, c.rep, c.state);
}

}  // namespace

// ======================================================================================= SA
Annealer::Annealer(const Hamiltonian &hamiltonian, AnnealSchedule schedule)
    : backend_(make_backend(BackendKind::HIP, hamiltonian)), schedule_(std::move(schedule)),
      rng_(std::random_device{}()) {
    if (schedule_.betas.empty()) throw std::invalid_argument("Schedule must contain at least one beta.");
}

Annealer::Annealer(std::shared_ptr<Backend> backend, AnnealSchedule schedule)
    : backend_(std::move(backend)), schedule_(std::move(schedule)), rng_(std::random_device{}()) {
    if (!backend_) throw std::invalid_argument("Annealer requires a backend.");
    if (schedule_.betas.empty()) throw std::invalid_argument("Schedule must contain at least one beta.");
}

void Annealer::set_seed(std::uint64_t seed) { rng_.seed(seed); }

AnnealResult Annealer::run(std::size_t sweeps_per_beta, Observer *observer) {
    if (sweeps_per_beta == 0) throw std::invalid_argument("sweeps_per_beta must be > 0.");
    HIPBackend &hb = device_backend(backend_);
    if (observer) return sa_observed(hb, schedule_, rng_, sweeps_per_beta, observer);
    const std::size_t n = hb.size(), steps = schedule_.size();
    qa_sa_params prm{};
    prm.steps = static_cast<int64_t>(steps);
    prm.betas = schedule_.betas.data();
    prm.sweeps_per_beta = static_cast<int64_t>(sweeps_per_beta);
    prm.reads = 1;
    prm.mode = mode_of(hb);
    State init;
    std::vector<double> uni;
    if (prm.mode == QA_MODE_FP32) {
        prm.seed = rng_();
    } else {
        init = State::random(n, rng_);  // annealer.cpp:40
        uni = uniform_stream(rng_, steps * sweeps_per_beta * n);
        prm.init_spins = init.spins.data();
        prm.uniforms = uni.data();
        prm.uniforms_per_read = static_cast<int64_t>(uni.size());
    }
    AnnealResult res;
    res.best_state = State(n);
    res.energy_trace.resize(steps);
    int64_t used = 0;
    qa_sa_result out{res.best_state.spins.data(), &res.best_energy, res.energy_trace.data(), &used, nullptr};
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sa_run(lease.ctx, lease.problem, &prm, &out));
    }
    if (prm.mode != QA_MODE_FP32) rng_.discard(static_cast<unsigned long long>(used));
    return res;
}

std::vector<AnnealResult> Annealer::run_reads(std::size_t sweeps_per_beta, std::size_t reads, std::uint64_t seed) {
    if (sweeps_per_beta == 0) throw std::invalid_argument("sweeps_per_beta must be > 0.");
    if (reads == 0) throw std::invalid_argument("reads must be > 0.");
    HIPBackend &hb = device_backend(backend_);
    const std::size_t n = hb.size(), steps = schedule_.size();
    qa_sa_params prm{};
    prm.steps = static_cast<int64_t>(steps);
    prm.betas = schedule_.betas.data();
    prm.sweeps_per_beta = static_cast<int64_t>(sweeps_per_beta);
    prm.reads = static_cast<int64_t>(reads);
    prm.mode = mode_of(hb);
    prm.seed = seed;
    std::vector<int8_t> init;
    std::vector<double> uni;
    if (prm.mode != QA_MODE_FP32) {
        const std::size_t per = steps * sweeps_per_beta * n;
        init.resize(reads * n);
        uni.resize(reads * per);
        for (std::size_t r = 0; r < reads; ++r) {
            std::mt19937_64 eng(seed + r);
            State s = State::random(n, eng);
            std::memcpy(&init[r * n], s.spins.data(), n);
            auto u = uniform_stream(eng, per);
            std::memcpy(&uni[r * per], u.data(), per * sizeof(double));
        }
        prm.init_spins = init.data();
        prm.uniforms = uni.data();
        prm.uniforms_per_read = static_cast<int64_t>(per);
    }
    std::vector<int8_t> best(reads * n);
    std::vector<double> be(reads), tr(reads * steps);
    qa_sa_result out{best.data(), be.data(), tr.data(), nullptr, nullptr};
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sa_run(lease.ctx, lease.problem, &prm, &out));
    }
    std::vector<AnnealResult> all(reads);
    for (std::size_t r = 0; r < reads; ++r) {
        all[r].best_state = state_from(&best[r * n], n);
        all[r].best_energy = be[r];
        all[r].energy_trace.assign(tr.begin() + r * steps, tr.begin() + (r + 1) * steps);
    }
    return all;
}

// ============================================================================ SA replicas
ReplicaAnnealer::ReplicaAnnealer(const Hamiltonian &hamiltonian, AnnealSchedule schedule, std::size_t replicas)
    : backend_(make_backend(BackendKind::HIP, hamiltonian)), schedule_(std::move(schedule)), replicas_(replicas),
      rng_(std::random_device{}()) {
    if (schedule_.betas.empty()) throw std::invalid_argument("Schedule must contain at least one beta.");
    if (replicas_ == 0) throw std::invalid_argument("replicas must be > 0.");
}

ReplicaAnnealer::ReplicaAnnealer(std::shared_ptr<Backend> backend, AnnealSchedule schedule, std::size_t replicas)
    : backend_(std::move(backend)), schedule_(std::move(schedule)), replicas_(replicas), rng_(std::random_device{}()) {
    if (!backend_) throw std::invalid_argument("ReplicaAnnealer requires a backend.");
    if (schedule_.betas.empty()) throw std::invalid_argument("Schedule must contain at least one beta.");
    if (replicas_ == 0) throw std::invalid_argument("replicas must be > 0.");
}

void ReplicaAnnealer::set_seed(std::uint64_t seed) { rng_.seed(seed); }

// replica_annealer.cpp:46-140. All replicas run as the reads of one SA launch. Validation mode
// draws what the reference draws from the annealer's engine: the R start states, then R replica
// seeds whose engines supply each replica's uniform stream. The global best is replayed on the
// host from each replica's best-so-far energy at the start and at every step end, with the
// reference's strict-less updates in replica order, so ties resolve to the same replica.
MultiAnnealResult ReplicaAnnealer::run(std::size_t sweeps_per_beta) {
    if (sweeps_per_beta == 0) throw std::invalid_argument("sweeps_per_beta must be > 0.");
    HIPBackend &hb = device_backend(backend_);
    const std::size_t n = hb.size(), steps = schedule_.size(), R = replicas_;
    qa_sa_params prm{};
    prm.steps = static_cast<int64_t>(steps);
    prm.betas = schedule_.betas.data();
    prm.sweeps_per_beta = static_cast<int64_t>(sweeps_per_beta);
    prm.reads = static_cast<int64_t>(R);
    prm.mode = mode_of(hb);
    std::vector<int8_t> init;
    std::vector<double> uni;
    if (prm.mode == QA_MODE_FP32) {
        prm.seed = rng_();
    } else {
        init.resize(R * n);
        for (std::size_t r = 0; r < R; ++r) {  // replica_annealer.cpp:56-59
            State st = State::random(n, rng_);
            std::memcpy(&init[r * n], st.spins.data(), n);
        }
        const std::size_t per = steps * sweeps_per_beta * n;
        uni.resize(R * per);
        for (std::size_t r = 0; r < R; ++r) {  // replica_annealer.cpp:60-63
            std::mt19937_64 eng(rng_());
            auto u = uniform_stream(eng, per);
            std::memcpy(&uni[r * per], u.data(), per * sizeof(double));
        }
        prm.init_spins = init.data();
        prm.uniforms = uni.data();
        prm.uniforms_per_read = static_cast<int64_t>(per);
    }
    std::vector<int8_t> best(R * n);
    std::vector<double> be(R), tr(R * steps), mg(R * steps), bt(R * (steps + 1));
    qa_sa_result out{};
    out.best_states = best.data();
    out.best_energies = be.data();
    out.energy_trace = tr.data();
    out.magnetization_trace = mg.data();
    out.best_energy_trace = bt.data();
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sa_run(lease.ctx, lease.problem, &prm, &out));
    }
    MultiAnnealResult res;
    res.replicas.resize(R);
    for (std::size_t r = 0; r < R; ++r) {
        res.replicas[r].best_state = state_from(&best[r * n], n);
        res.replicas[r].best_energy = be[r];
        res.replicas[r].energy_trace.assign(tr.begin() + r * steps, tr.begin() + (r + 1) * steps);
        res.replicas[r].magnetization_trace.assign(mg.begin() + r * steps, mg.begin() + (r + 1) * steps);
    }
    double g = std::numeric_limits<double>::infinity();
    std::size_t owner = R;
    for (std::size_t k = 0; k <= steps; ++k)  // k = 0: start states (replica_annealer.cpp:77-80)
        for (std::size_t r = 0; r < R; ++r)
            if (bt[r * (steps + 1) + k] < g) {
                g = bt[r * (steps + 1) + k];
                owner = r;
            }
    res.global_best_energy = owner < R ? be[owner] : g;
    if (owner < R) res.global_best_state = res.replicas[owner].best_state;
    res.average_energy_trace.resize(steps);
    res.average_magnetization_trace.resize(steps);
    for (std::size_t k = 0; k < steps; ++k) {  // replica_annealer.cpp:118-133, replica order
        double ae = 0.0, am = 0.0;
        for (std::size_t r = 0; r < R; ++r) {
            ae += tr[r * steps + k];
            am += mg[r * steps + k];
        }
        res.average_energy_trace[k] = ae / static_cast<double>(R);
        res.average_magnetization_trace[k] = am / static_cast<double>(R);
    }
    return res;
}

// ==================================================================== parallel tempering
ParallelTemperingAnnealer::ParallelTemperingAnnealer(const Hamiltonian &hamiltonian, std::vector<double> betas)
    : backend_(make_backend(BackendKind::HIP, hamiltonian)), betas_(std::move(betas)), rng_(std::random_device{}()) {
    if (betas_.size() < 2) throw std::invalid_argument("Parallel tempering requires at least two betas.");
}

ParallelTemperingAnnealer::ParallelTemperingAnnealer(std::shared_ptr<Backend> backend, std::vector<double> betas)
    : backend_(std::move(backend)), betas_(std::move(betas)), rng_(std::random_device{}()) {
    if (!backend_) throw std::invalid_argument("ParallelTemperingAnnealer requires a backend.");
    if (betas_.size() < 2) throw std::invalid_argument("Parallel tempering requires at least two betas.");
}

void ParallelTemperingAnnealer::set_seed(std::uint64_t seed) { rng_.seed(seed); }

// parallel_tempering.cpp:36-122. FP32: each step is one SA launch over all rungs (read r at
// betas[r]) continuing from the current states and exact energies, with the step index as the
// Philox stream offset; the final energies come back recomputed in FP64, so swaps and the best
// state see exact energies. Validation: rung r's sweeps of a step consume the annealer's engine
// right after rung r-1's, so each rung-step is its own launch on the engine's current stream,
// and the engine is advanced by what the device used. Swaps draw from the same engine.
ParallelTemperingResult ParallelTemperingAnnealer::run(std::size_t sweeps_per_step, std::size_t steps,
                                                       std::size_t swap_interval) {
    if (sweeps_per_step == 0) throw std::invalid_argument("sweeps_per_step must be > 0.");
    if (steps == 0) throw std::invalid_argument("steps must be > 0.");
    if (swap_interval == 0) throw std::invalid_argument("swap_interval must be > 0.");
    HIPBackend &hb = device_backend(backend_);
    const std::size_t n = hb.size(), R = betas_.size();
    const int mode = mode_of(hb);
    std::vector<int8_t> cur(R * n), nxt(R * n), best(R * n);
    std::vector<double> energies(R), fe(R), be(R), tr(R);
    for (std::size_t r = 0; r < R; ++r) {  // parallel_tempering.cpp:54-57
        State st = State::random(n, rng_);
        std::memcpy(&cur[r * n], st.spins.data(), n);
    }
    const uint64_t key = mode == QA_MODE_FP32 ? rng_() : 0;
    ParallelTemperingResult res;
    res.best_energy = std::numeric_limits<double>::infinity();
    res.average_energy_trace.reserve(steps);
    res.swap_acceptance_trace.reserve(steps);
    std::uniform_real_distribution<double> uniform(0.0, 1.0);
    const auto lease = hb.lease();
    for (std::size_t step = 0; step < steps; ++step) {
        qa_sa_params prm{};
        prm.steps = 1;
        prm.sweeps_per_beta = static_cast<int64_t>(sweeps_per_step);
        prm.mode = mode;
        qa_sa_result out{};
        if (mode == QA_MODE_FP32) {
            prm.betas = betas_.data();
            prm.read_betas = betas_.data();
            prm.reads = static_cast<int64_t>(R);
            prm.seed = key;
            prm.init_states = cur.data();
            prm.init_energies = step == 0 ? nullptr : energies.data();
            prm.counter_base = step;
            out.best_states = best.data();
            out.best_energies = be.data();
            out.energy_trace = tr.data();
            out.final_states = nxt.data();
            out.final_energies = fe.data();
            throw_on_status(qa_sa_run(lease.ctx, lease.problem, &prm, &out));
        } else {
            for (std::size_t r = 0; r < R; ++r) {
                const auto uni = uniform_stream(rng_, sweeps_per_step * n);
                int64_t used = 0;
                prm.betas = &betas_[r];
                prm.reads = 1;
                prm.init_spins = &cur[r * n];
                prm.init_energies = step == 0 ? nullptr : &energies[r];
                prm.uniforms = uni.data();
                prm.uniforms_per_read = static_cast<int64_t>(uni.size());
                out.best_states = &best[r * n];
                out.best_energies = &be[r];
                out.energy_trace = &tr[r];
                out.uniforms_used = &used;
                out.final_states = &nxt[r * n];
                out.final_energies = &fe[r];
                throw_on_status(qa_sa_run(lease.ctx, lease.problem, &prm, &out));
                rng_.discard(static_cast<unsigned long long>(used));
            }
        }
        cur.swap(nxt);
        energies = fe;
        for (std::size_t r = 0; r < R; ++r)  // parallel_tempering.cpp:84-87
            if (energies[r] < res.best_energy) {
                res.best_energy = energies[r];
                res.best_state = state_from(&cur[r * n], n);
            }
        double accepted = 0.0, attempted = 0.0;
        if ((step + 1) % swap_interval == 0) {  // parallel_tempering.cpp:92-106
            for (std::size_t r = 0; r + 1 < R; ++r) {
                const double delta = (betas_[r] - betas_[r + 1]) * (energies[r + 1] - energies[r]);
                ++attempted;
                if (delta <= 0.0 || uniform(rng_) < std::exp(-delta)) {
                    std::swap_ranges(cur.begin() + r * n, cur.begin() + (r + 1) * n, cur.begin() + (r + 1) * n);
                    std::swap(energies[r], energies[r + 1]);
                    ++accepted;
                }
            }
        }
        double avg = 0.0;
        for (double e : energies) avg += e;
        res.average_energy_trace.push_back(avg / static_cast<double>(R));
        res.swap_acceptance_trace.push_back(attempted > 0.0 ? accepted / attempted : 0.0);
    }
    res.final_states.reserve(R);
    for (std::size_t r = 0; r < R; ++r) res.final_states.push_back(state_from(&cur[r * n], n));
    res.final_energies = energies;
    return res;
}

// ====================================================================================== SQA
namespace {
void check_sqa_ctor(const SQASchedule &s, std::size_t slices, std::size_t replicas) {
    if (s.betas.empty()) throw std::invalid_argument("SQA schedule must contain betas.");
    if (s.betas.size() != s.gammas.size()) throw std::invalid_argument("SQA schedule betas/gammas length mismatch.");
    if (slices == 0) throw std::invalid_argument("trotter_slices must be > 0.");
    if (replicas == 0) throw std::invalid_argument("replicas must be > 0.");
}
}  // namespace

SQAAnnealer::SQAAnnealer(const Hamiltonian &hamiltonian, SQASchedule schedule, std::size_t trotter_slices,
                         std::size_t replicas)
    : backend_(make_backend(BackendKind::HIP, hamiltonian)), schedule_(std::move(schedule)),
      slices_(trotter_slices), replicas_(replicas), rng_(std::random_device{}()) {
    check_sqa_ctor(schedule_, slices_, replicas_);
}

SQAAnnealer::SQAAnnealer(std::shared_ptr<Backend> backend, SQASchedule schedule, std::size_t trotter_slices,
                         std::size_t replicas)
    : backend_(std::move(backend)), schedule_(std::move(schedule)), slices_(trotter_slices),
      replicas_(replicas), rng_(std::random_device{}()) {
    if (!backend_) throw std::invalid_argument("SQAAnnealer requires a backend.");
    check_sqa_ctor(schedule_, slices_, replicas_);
}

void SQAAnnealer::set_seed(std::uint64_t seed) { rng_.seed(seed); }

SQAResult SQAAnnealer::run(std::size_t sweeps_per_beta, std::size_t worldline_sweeps, std::size_t cluster_sweeps,
                           std::size_t continuous_time_slices, SQAObserver *observer) {
    if (sweeps_per_beta == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
        throw std::invalid_argument(
            "At least one of sweeps_per_beta, worldline_sweeps, or cluster_sweeps must be > 0.");
    HIPBackend &hb = device_backend(backend_);
    const std::size_t n = hb.size();
    const std::size_t P = continuous_time_slices > slices_ ? continuous_time_slices : slices_;  // :97-99
    const std::size_t R = replicas_, steps = schedule_.size();
    qa_sqa_params prm{};
    prm.steps = static_cast<int64_t>(steps);
    prm.betas = schedule_.betas.data();
    prm.gammas = schedule_.gammas.data();
    prm.trotter_slices = static_cast<int64_t>(P);
    prm.replicas = static_cast<int64_t>(R);
    prm.sweeps_per_beta = static_cast<int64_t>(sweeps_per_beta);
    prm.worldline_sweeps = static_cast<int64_t>(worldline_sweeps);
    prm.cluster_sweeps = static_cast<int64_t>(cluster_sweeps);
    prm.mode = mode_of(hb);
    SQAState init;
    std::vector<uint64_t> seeds;
    auto *sweep_obs = dynamic_cast<SQASweepObserver *>(observer);
    std::string engine_text;
    if (prm.mode == QA_MODE_FP32 && !sweep_obs) {
        prm.seed = rng_();
    } else {
        init = SQAState::random(R, P, n, rng_);  // :100
        seeds.resize(R);
        for (auto &s : seeds) s = rng_();        // :108-111
        prm.init_spins = init.data();
        prm.replica_seeds = seeds.data();
    }
    SqaObserverCtx octx{observer, sweep_obs, &schedule_, observer ? SQAState(R, P, n) : SQAState(), {}};
    if (observer) {
        prm.on_step = sqa_on_step;
        prm.callback_user = &octx;
    }
    std::vector<char> engine_out;
    if (sweep_obs) {  // the tracing path continues the annealer's engine (:317-453)
        std::ostringstream o;
        o << rng_;
        engine_text = o.str();
        prm.on_sweep = sqa_on_sweep;
        prm.trace_engine = engine_text.c_str();
        engine_out.assign(engine_text.size() + 4096, '\0');
    }
    SQAResult res;
    res.best_state = State(n);
    res.energy_trace.resize(steps);
    double kms = 0.0, acc = 0.0;
    qa_sqa_result out{res.best_state.spins.data(), &res.best_energy, res.energy_trace.data(), &kms, &acc};
    if (sweep_obs) {
        out.trace_engine_out = engine_out.data();
        out.trace_engine_out_len = static_cast<int64_t>(engine_out.size());
    }
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sqa_run(lease.ctx, lease.problem, &prm, &out));
    }
    if (sweep_obs) {
        std::istringstream in(engine_out.data());
        in >> rng_;
    }
    last_kernel_ms_ = kms;
    last_accept_rate_ = acc;
    return res;
}

// sqa_annealer.cpp:499-952. Validation draws the start state and replica seeds as run() does
// (:534-541); the pilot's random draws come from the replicas' own engines.
SQAResult SQAAnnealer::run_optimal(double beta, double j_perp_start, double j_perp_end, double eps_tilde, double alpha,
                                   std::size_t num_steps, std::size_t sweeps_per_step, std::size_t worldline_sweeps,
                                   std::size_t cluster_sweeps, std::size_t calib_probes, std::size_t calib_sweeps,
                                   const std::string &debug_csv_path, double beta_ramp_fraction,
                                   double beta_ramp_start) {
    if (num_steps == 0) throw std::invalid_argument("num_steps must be > 0.");
    if (sweeps_per_step == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
        throw std::invalid_argument(
            "At least one of sweeps_per_step, worldline_sweeps, or cluster_sweeps must be > 0.");
    if (j_perp_end <= j_perp_start) throw std::invalid_argument("j_perp_end must be > j_perp_start.");
    HIPBackend &hb = device_backend(backend_);
    const std::size_t n = hb.size(), P = slices_, R = replicas_;
    qa_sqa_opt_params prm{};
    prm.beta = beta;
    prm.j_perp_start = j_perp_start;
    prm.j_perp_end = j_perp_end;
    prm.eps_tilde = eps_tilde;
    prm.alpha = alpha;
    prm.num_steps = static_cast<int64_t>(num_steps);
    prm.sweeps_per_step = static_cast<int64_t>(sweeps_per_step);
    prm.worldline_sweeps = static_cast<int64_t>(worldline_sweeps);
    prm.cluster_sweeps = static_cast<int64_t>(cluster_sweeps);
    prm.calib_probes = static_cast<int64_t>(calib_probes);
    prm.calib_sweeps = static_cast<int64_t>(calib_sweeps);
    prm.beta_ramp_fraction = beta_ramp_fraction;
    prm.beta_ramp_start = beta_ramp_start;
    prm.trotter_slices = static_cast<int64_t>(P);
    prm.replicas = static_cast<int64_t>(R);
    prm.mode = mode_of(hb);
    prm.debug_csv_path = debug_csv_path.c_str();
    SQAState init;
    std::vector<uint64_t> seeds;
    if (prm.mode == QA_MODE_FP32) {
        prm.seed = rng_();
    } else {
        init = SQAState::random(R, P, n, rng_);
        seeds.resize(R);
        for (auto &s : seeds) s = rng_();
        prm.init_spins = init.data();
        prm.replica_seeds = seeds.data();
    }
    SQAResult res;
    res.best_state = State(n);
    res.energy_trace.resize(num_steps);
    res.j_perp_trace.resize(num_steps);
    res.chi_B_trace.resize(num_steps);
    double kms = 0.0;
    qa_sqa_opt_result out{};
    out.best_state = res.best_state.spins.data();
    out.best_energy = &res.best_energy;
    out.energy_trace = res.energy_trace.data();
    out.sched = qa_opt_schedule{res.j_perp_trace.data(), res.chi_B_trace.data(), &res.calibrated_eps_tilde,
                                &res.j_perp_start, &res.resolved_j_perp_end, &res.final_j_perp};
    out.kernel_ms = &kms;
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sqa_optimal_run(lease.ctx, lease.problem, &prm, &out));
    }
    last_kernel_ms_ = kms;
    last_accept_rate_ = 0.0;
    return res;
}

// ================================================================================== SQA-PT
namespace {
void check_pt_ctor(const std::vector<double> &b, const std::vector<double> &g, std::size_t slices) {
    if (b.size() < 2 || b.size() != g.size())
        throw std::invalid_argument("SQA parallel tempering requires >=2 beta/gamma pairs.");
    if (slices == 0) throw std::invalid_argument("trotter_slices must be > 0.");
}
}  // namespace

SQAParallelTemperingAnnealer::SQAParallelTemperingAnnealer(const Hamiltonian &hamiltonian, std::vector<double> betas,
                                                           std::vector<double> gammas, std::size_t trotter_slices)
    : backend_(make_backend(BackendKind::HIP, hamiltonian)), betas_(std::move(betas)), gammas_(std::move(gammas)),
      slices_(trotter_slices), rng_(std::random_device{}()) {
    check_pt_ctor(betas_, gammas_, slices_);
}

SQAParallelTemperingAnnealer::SQAParallelTemperingAnnealer(std::shared_ptr<Backend> backend,
                                                           std::vector<double> betas, std::vector<double> gammas,
                                                           std::size_t trotter_slices)
    : backend_(std::move(backend)), betas_(std::move(betas)), gammas_(std::move(gammas)), slices_(trotter_slices),
      rng_(std::random_device{}()) {
    if (!backend_) throw std::invalid_argument("SQAParallelTemperingAnnealer requires a backend.");
    check_pt_ctor(betas_, gammas_, slices_);
}

void SQAParallelTemperingAnnealer::set_seed(std::uint64_t seed) { rng_.seed(seed); }

SQAParallelTemperingResult SQAParallelTemperingAnnealer::run(std::size_t sweeps_per_step,
                                                             std::size_t worldline_sweeps, std::size_t steps,
                                                             std::size_t swap_interval, std::size_t cluster_sweeps,
                                                             std::size_t continuous_time_slices) {
    if (steps == 0) throw std::invalid_argument("steps must be > 0.");
    if (swap_interval == 0) throw std::invalid_argument("swap_interval must be > 0.");
    if (sweeps_per_step == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
        throw std::invalid_argument("At least one of sweeps_per_step/worldline_sweeps/cluster_sweeps must be > 0.");
    HIPBackend &hb = device_backend(backend_);
    const std::size_t n = hb.size(), R = betas_.size();
    const std::size_t P = continuous_time_slices > slices_ ? continuous_time_slices : slices_;
    qa_sqapt_params prm{};
    prm.rungs = static_cast<int64_t>(R);
    prm.betas = betas_.data();
    prm.gammas = gammas_.data();
    prm.ladders = 1;
    prm.trotter_slices = static_cast<int64_t>(P);
    prm.sweeps_per_step = static_cast<int64_t>(sweeps_per_step);
    prm.worldline_sweeps = static_cast<int64_t>(worldline_sweeps);
    prm.cluster_sweeps = static_cast<int64_t>(cluster_sweeps);
    prm.steps = static_cast<int64_t>(steps);
    prm.swap_interval = static_cast<int64_t>(swap_interval);
    prm.mode = mode_of(hb);
    std::vector<int8_t> init;
    std::vector<uint64_t> seeds;
    std::vector<double> swap_u;
    if (prm.mode == QA_MODE_FP32) {
        prm.seed = rng_();
    } else {
        init.resize(R * P * n);
        for (std::size_t r = 0; r < R; ++r) {  // sqa_parallel_tempering.cpp:103-107
            SQAState s = SQAState::random(1, P, n, rng_);
            std::memcpy(&init[r * P * n], s.data(), P * n);
        }
        seeds.resize(R);
        for (auto &s : seeds) s = rng_();  // :109-112
        swap_u = uniform_stream(rng_, steps * (R - 1));
        prm.init_spins = init.data();
        prm.rung_seeds = seeds.data();
        prm.swap_uniforms = swap_u.data();
        prm.swap_uniforms_len = static_cast<int64_t>(swap_u.size());
    }
    std::vector<int8_t> fs(R * n), bs(n);
    std::vector<double> fe(R), at(steps), st(steps);
    double be = 0.0;
    int64_t used = 0;
    qa_sqapt_result out{fs.data(), fe.data(), bs.data(), &be, at.data(), st.data(), &used, nullptr};
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sqapt_run(lease.ctx, lease.problem, &prm, &out));
    }
    if (prm.mode != QA_MODE_FP32) rng_.discard(static_cast<unsigned long long>(used));
    SQAParallelTemperingResult res;
    for (std::size_t r = 0; r < R; ++r) res.final_states.push_back(state_from(&fs[r * n], n));
    res.final_energies = fe;
    res.best_state = state_from(bs.data(), n);
    res.best_energy = be;
    res.average_energy_trace = at;
    res.swap_acceptance_trace = st;
    return res;
}

std::vector<SQAParallelTemperingResult> SQAParallelTemperingAnnealer::run_ladders(
    std::size_t ladders, std::size_t sweeps_per_step, std::size_t worldline_sweeps, std::size_t steps,
    std::size_t swap_interval, std::size_t cluster_sweeps, const std::vector<std::vector<double>> &step_betas,
    const std::vector<std::vector<double>> &step_gammas) {
    if (ladders == 0) throw std::invalid_argument("ladders must be > 0.");
    if (steps == 0) throw std::invalid_argument("steps must be > 0.");
    if (swap_interval == 0) throw std::invalid_argument("swap_interval must be > 0.");
    HIPBackend &hb = device_backend(backend_);
    if (hb.precision() != DevicePrecision::FP32)
        throw std::invalid_argument("run_ladders is an FP32-mode entry point");
    const std::size_t n = hb.size(), R = betas_.size(), L = ladders;
    if (step_betas.empty() != step_gammas.empty())
        throw std::invalid_argument("step_betas and step_gammas go together");
    std::vector<double> sb, sg;  // [steps*R] row-major
    if (!step_betas.empty()) {
        if (step_betas.size() != steps || step_gammas.size() != steps)
            throw std::invalid_argument("step_betas / step_gammas need one row per step");
        for (std::size_t k = 0; k < steps; ++k) {
            if (step_betas[k].size() != R || step_gammas[k].size() != R)
                throw std::invalid_argument("every step_betas / step_gammas row needs one value per rung");
            sb.insert(sb.end(), step_betas[k].begin(), step_betas[k].end());
            sg.insert(sg.end(), step_gammas[k].begin(), step_gammas[k].end());
        }
    }
    qa_sqapt_params prm{};
    prm.rungs = static_cast<int64_t>(R);
    prm.betas = betas_.data();
    prm.gammas = gammas_.data();
    prm.ladders = static_cast<int64_t>(L);
    prm.trotter_slices = static_cast<int64_t>(slices_);
    prm.sweeps_per_step = static_cast<int64_t>(sweeps_per_step);
    prm.worldline_sweeps = static_cast<int64_t>(worldline_sweeps);
    prm.cluster_sweeps = static_cast<int64_t>(cluster_sweeps);
    prm.steps = static_cast<int64_t>(steps);
    prm.swap_interval = static_cast<int64_t>(swap_interval);
    prm.mode = QA_MODE_FP32;
    prm.seed = rng_();
    if (!sb.empty()) {
        prm.step_betas = sb.data();
        prm.step_gammas = sg.data();
    }
    std::vector<int8_t> fs(L * R * n), bs(L * n);
    std::vector<double> fe(L * R), be(L), at(L * steps), st(L * steps);
    qa_sqapt_result out{fs.data(), fe.data(), bs.data(), be.data(), at.data(), st.data(), nullptr, nullptr};
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sqapt_run(lease.ctx, lease.problem, &prm, &out));
    }
    std::vector<SQAParallelTemperingResult> all(L);
    for (std::size_t l = 0; l < L; ++l) {
        auto &r = all[l];
        for (std::size_t q = 0; q < R; ++q) r.final_states.push_back(state_from(&fs[(l * R + q) * n], n));
        r.final_energies.assign(fe.begin() + l * R, fe.begin() + (l + 1) * R);
        r.best_state = state_from(&bs[l * n], n);
        r.best_energy = be[l];
        r.average_energy_trace.assign(at.begin() + l * steps, at.begin() + (l + 1) * steps);
        r.swap_acceptance_trace.assign(st.begin() + l * steps, st.begin() + (l + 1) * steps);
    }
    return all;
}

// sqa_parallel_tempering.cpp:348-882. Validation draws, in the reference's order: the rungs'
// start states (:406-409), the rung seeds (:411-412), the pilot's scratch seeds when calibrating
// (:655-656); swaps then consume the engine's uniforms, and the engine is advanced by them.
SQAParallelTemperingResult SQAParallelTemperingAnnealer::run_optimal(
    std::size_t num_steps, std::size_t sweeps_per_step, std::size_t worldline_sweeps, double eps_tilde, double alpha,
    double j_perp_end, std::size_t cluster_sweeps, std::size_t swap_interval, std::size_t continuous_time_slices,
    std::size_t calib_probes, std::size_t calib_sweeps, const std::string &debug_csv_path) {
    if (num_steps == 0) throw std::invalid_argument("num_steps must be > 0.");
    if (swap_interval == 0) throw std::invalid_argument("swap_interval must be > 0.");
    if (sweeps_per_step == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
        throw std::invalid_argument("At least one of sweeps_per_step/worldline_sweeps/cluster_sweeps must be > 0.");
    HIPBackend &hb = device_backend(backend_);
    const std::size_t n = hb.size(), R = betas_.size();
    const std::size_t P =
        (continuous_time_slices > 0 && continuous_time_slices > slices_) ? continuous_time_slices : slices_;
    qa_sqapt_opt_params prm{};
    prm.rungs = static_cast<int64_t>(R);
    prm.betas = betas_.data();
    prm.gammas = gammas_.data();
    prm.trotter_slices = static_cast<int64_t>(P);
    prm.num_steps = static_cast<int64_t>(num_steps);
    prm.sweeps_per_step = static_cast<int64_t>(sweeps_per_step);
    prm.worldline_sweeps = static_cast<int64_t>(worldline_sweeps);
    prm.cluster_sweeps = static_cast<int64_t>(cluster_sweeps);
    prm.swap_interval = static_cast<int64_t>(swap_interval);
    prm.calib_probes = static_cast<int64_t>(calib_probes);
    prm.calib_sweeps = static_cast<int64_t>(calib_sweeps);
    prm.eps_tilde = eps_tilde;
    prm.alpha = alpha;
    prm.j_perp_end = j_perp_end;
    prm.mode = mode_of(hb);
    prm.debug_csv_path = debug_csv_path.c_str();
    std::vector<int8_t> init;
    std::vector<uint64_t> seeds, scratch;
    std::vector<double> swap_u;
    if (prm.mode == QA_MODE_FP32) {
        prm.seed = rng_();
    } else {
        init.resize(R * P * n);
        for (std::size_t r = 0; r < R; ++r) {
            SQAState s = SQAState::random(1, P, n, rng_);
            std::memcpy(&init[r * P * n], s.data(), P * n);
        }
        seeds.resize(R);
        for (auto &s : seeds) s = rng_();
        if (eps_tilde <= 0.0) {
            scratch.resize(R);
            for (auto &s : scratch) s = rng_();
        }
        swap_u = uniform_stream(rng_, (num_steps / swap_interval) * (R - 1));
        prm.init_spins = init.data();
        prm.rung_seeds = seeds.data();
        prm.scratch_seeds = scratch.empty() ? nullptr : scratch.data();
        prm.swap_uniforms = swap_u.data();
        prm.swap_uniforms_len = static_cast<int64_t>(swap_u.size());
    }
    SQAParallelTemperingResult res;
    std::vector<int8_t> fs(R * n), bs(n);
    res.final_energies.resize(R);
    res.average_energy_trace.resize(num_steps);
    res.swap_acceptance_trace.resize(num_steps);
    res.j_perp_trace.resize(num_steps);
    res.chi_B_trace.resize(num_steps);
    int64_t used = 0;
    qa_sqapt_opt_result out{};
    out.final_states = fs.data();
    out.final_energies = res.final_energies.data();
    out.best_state = bs.data();
    out.best_energy = &res.best_energy;
    out.average_energy_trace = res.average_energy_trace.data();
    out.swap_acceptance_trace = res.swap_acceptance_trace.data();
    out.sched = qa_opt_schedule{res.j_perp_trace.data(), res.chi_B_trace.data(), &res.calibrated_eps_tilde,
                                &res.j_perp_start, &res.resolved_j_perp_end, &res.final_j_perp};
    out.swap_uniforms_used = &used;
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sqapt_optimal_run(lease.ctx, lease.problem, &prm, &out));
    }
    if (prm.mode != QA_MODE_FP32) rng_.discard(static_cast<unsigned long long>(used));
    for (std::size_t r = 0; r < R; ++r) res.final_states.push_back(state_from(&fs[r * n], n));
    res.best_state = state_from(bs.data(), n);
    return res;
}

// ================================================================================= sqa_chi
SQAChiAnnealer::SQAChiAnnealer(const Hamiltonian &hamiltonian, std::size_t trotter_slices, std::size_t replicas)
    : backend_(make_backend(BackendKind::HIP, hamiltonian)), slices_(trotter_slices), replicas_(replicas),
      rng_(std::random_device{}()) {
    if (slices_ == 0) throw std::invalid_argument("trotter_slices must be > 0.");
    if (replicas_ == 0) throw std::invalid_argument("replicas must be > 0.");
}

SQAChiAnnealer::SQAChiAnnealer(std::shared_ptr<Backend> backend, std::size_t trotter_slices, std::size_t replicas)
    : backend_(std::move(backend)), slices_(trotter_slices), replicas_(replicas), rng_(std::random_device{}()) {
    if (!backend_) throw std::invalid_argument("SQAChiAnnealer requires a backend.");
    if (slices_ == 0) throw std::invalid_argument("trotter_slices must be > 0.");
    if (replicas_ == 0) throw std::invalid_argument("replicas must be > 0.");
}

void SQAChiAnnealer::set_seed(std::uint64_t seed) { rng_.seed(seed); }

namespace {
SQAChiResult chi_run(HIPBackend &hb, std::mt19937_64 &rng, std::size_t P, std::size_t R, qa_chi_params prm) {
    const std::size_t n = hb.size();
    const std::size_t NS = static_cast<std::size_t>(prm.num_steps), SP = static_cast<std::size_t>(prm.scan_points);
    std::vector<int8_t> init;
    std::vector<uint64_t> seeds;
    if (prm.mode == QA_MODE_FP32) {
        prm.seed = rng();
    } else {
        // every fresh worldline draws R*P*n spins then R*P cell seeds (sqa_chi_annealer.cpp:228-240,
        // 394-405): scan_points pilot worldlines, then the production one
        const std::size_t blocks = SP + 1;
        init.resize(blocks * R * P * n);
        seeds.resize(blocks * R * P);
        for (std::size_t b = 0; b < blocks; ++b) {
            SQAState s = SQAState::random(R, P, n, rng);
            std::memcpy(&init[b * R * P * n], s.data(), R * P * n);
            for (std::size_t k = 0; k < R * P; ++k) seeds[b * R * P + k] = rng();
        }
        prm.init_spins = init.data();
        prm.cell_seeds = seeds.data();
    }
    SQAChiResult res;
    res.best_state = State(n);
    res.energy_trace.resize(NS);
    res.beta_schedule.resize(NS);
    res.gamma_schedule.resize(NS);
    res.s_schedule.resize(NS);
    res.scan_s.resize(SP);
    res.scan_gamma.resize(SP);
    res.scan_chi_B.resize(SP);
    double sc[5] = {0, 0, 0, 0, 0};
    qa_chi_result out{res.best_state.spins.data(), &res.best_energy, res.energy_trace.data(),
                      res.beta_schedule.data(), res.gamma_schedule.data(), res.s_schedule.data(),
                      res.scan_s.data(), res.scan_gamma.data(), res.scan_chi_B.data(), sc, nullptr};
    {
        const auto lease = hb.lease();
        throw_on_status(qa_sqa_chi_run(lease.ctx, lease.problem, &prm, &out));
    }
    res.s_star = sc[0];
    res.gamma_star = sc[1];
    res.j_perp_star = sc[2];
    res.chi_floor = sc[3];
    res.driver_A0 = sc[4];
    return res;
}
}  // namespace

SQAChiResult SQAChiAnnealer::run_chi(double beta, double gamma_start, double gamma_end, std::size_t num_steps,
                                     std::size_t sweeps_per_step, std::size_t scan_points, std::size_t scan_sweeps,
                                     std::size_t scan_burn, double chi_floor_fraction, double driver_A0,
                                     double beta_ramp_fraction, double beta_ramp_start, const std::string &) {
    // sqa_chi_annealer.cpp:159-179 argument checks
    if (num_steps < 2) throw std::invalid_argument("num_steps must be >= 2.");
    if (sweeps_per_step == 0) throw std::invalid_argument("sweeps_per_step must be > 0.");
    if (!(beta > 0.0)) throw std::invalid_argument("beta must be > 0.");
    if (!(gamma_end > 0.0) || !(gamma_start > gamma_end))
        throw std::invalid_argument("Require gamma_start > gamma_end > 0.");
    if (scan_points < 4) throw std::invalid_argument("scan_points must be >= 4.");
    if (scan_sweeps == 0) throw std::invalid_argument("scan_sweeps must be > 0.");
    if (chi_floor_fraction < 0.0) throw std::invalid_argument("chi_floor_fraction must be >= 0.");
    HIPBackend &hb = device_backend(backend_);
    qa_chi_params prm{};
    prm.trotter_slices = static_cast<int64_t>(slices_);
    prm.replicas = static_cast<int64_t>(replicas_);
    prm.beta = beta;
    prm.gamma_start = gamma_start;
    prm.gamma_end = gamma_end;
    prm.num_steps = static_cast<int64_t>(num_steps);
    prm.sweeps_per_step = static_cast<int64_t>(sweeps_per_step);
    prm.scan_points = static_cast<int64_t>(scan_points);
    prm.scan_sweeps = static_cast<int64_t>(scan_sweeps);
    prm.scan_burn = static_cast<int64_t>(scan_burn);
    prm.chi_floor_fraction = chi_floor_fraction;
    prm.driver_A0 = driver_A0;
    prm.beta_ramp_fraction = beta_ramp_fraction;
    prm.beta_ramp_start = beta_ramp_start;
    prm.mode = mode_of(hb);
    return chi_run(hb, rng_, slices_, replicas_, prm);
}

SQAChiResult SQAChiAnnealer::run_chi_with_profile(const std::vector<double> &chi_B, double beta, double gamma_start,
                                                  double gamma_end, std::size_t num_steps,
                                                  std::size_t sweeps_per_step, double chi_floor_fraction,
                                                  double driver_A0, double beta_ramp_fraction,
                                                  double beta_ramp_start) {
    if (chi_B.size() < 4) throw std::invalid_argument("scan_points must be >= 4.");
    if (num_steps < 2) throw std::invalid_argument("num_steps must be >= 2.");
    if (sweeps_per_step == 0) throw std::invalid_argument("sweeps_per_step must be > 0.");
    if (!(beta > 0.0)) throw std::invalid_argument("beta must be > 0.");
    if (!(gamma_end > 0.0) || !(gamma_start > gamma_end))
        throw std::invalid_argument("Require gamma_start > gamma_end > 0.");
    HIPBackend &hb = device_backend(backend_);
    if (hb.precision() != DevicePrecision::FP32)
        throw std::invalid_argument("a supplied chi_B profile is an FP32-mode path");
    qa_chi_params prm{};
    prm.trotter_slices = static_cast<int64_t>(slices_);
    prm.replicas = static_cast<int64_t>(replicas_);
    prm.beta = beta;
    prm.gamma_start = gamma_start;
    prm.gamma_end = gamma_end;
    prm.num_steps = static_cast<int64_t>(num_steps);
    prm.sweeps_per_step = static_cast<int64_t>(sweeps_per_step);
    prm.scan_points = static_cast<int64_t>(chi_B.size());
    prm.scan_sweeps = 1;
    prm.scan_burn = 0;
    prm.chi_floor_fraction = chi_floor_fraction;
    prm.driver_A0 = driver_A0;
    prm.beta_ramp_fraction = beta_ramp_fraction;
    prm.beta_ramp_start = beta_ramp_start;
    prm.mode = QA_MODE_FP32;
    prm.chi_override = chi_B.data();
    return chi_run(hb, rng_, slices_, replicas_, prm);
}

}  // namespace qanneal
