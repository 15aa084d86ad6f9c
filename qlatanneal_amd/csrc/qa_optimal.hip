// Adaptive transverse-coupling schedules (SQAAnnealer::run_optimal, sqa_annealer.cpp:499-952;
// SQAParallelTemperingAnnealer::run_optimal, sqa_parallel_tempering.cpp:348-882).
//
// The schedule logic runs on the host, as in the reference; the sweeps run on the device through
// two step engines: FP32 (SqaFp32Session: multi-step launches that keep the state on the device,
// with per-step bond statistics) and FP64 validation (SqaValSession / PtValSession: the reference
// trajectory replayed phase by phase). chi_B = Var(B) is formed from the exact integer sums of B
// and B^2, so both engines feed the schedule the reference's own values. The host arithmetic
// writes the reference build's floating-point contraction sites as explicit fma (everything else
// uncontracted), so schedules are bit-identical to the reference's for the same chi_B.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <random>
#include <vector>

#include "qa_internal.hpp"

namespace qa {
namespace {

// chi_B profile from the calibration probes: linear between probes, clamped to the probe range;
// -1 without a profile (sqa_annealer.cpp:880-891).
struct ChiProfile {
    std::vector<double> J, chi;
    double at(double x) const {
        const size_t M = J.size();
        if (M == 0) return -1.0;
        if (x <= J.front()) return chi.front();
        if (x >= J.back()) return chi.back();
        size_t hi = 1;
        while (hi < M && J[hi] < x) ++hi;
        const size_t lo = hi - 1;
        const double t = (x - J[lo]) / (J[hi] - J[lo]);
        return std::fma(t, chi[hi] - chi[lo], chi[lo]);
    }
};

// J_perp(step) by inverting G(J) = integral of max(chi_B(J), floor)^alpha over a 4096-cell
// trapezoid grid, endpoints inclusive (sqa_annealer.cpp:895-921).
std::vector<double> inverse_cdf_trajectory(const ChiProfile &pr, double j0, double j1, double floor_chi,
                                           double alpha, size_t count) {
    const size_t GRID = 4096;
    const double h = (j1 - j0) / static_cast<double>(GRID);
    std::vector<double> G(GRID + 1, 0.0);
    double prev = std::pow(std::max(pr.at(j0), floor_chi), alpha);
    for (size_t k = 1; k <= GRID; ++k) {
        const double Jk = std::fma(static_cast<double>(k), h, j0);
        const double cur = std::pow(std::max(pr.at(Jk), floor_chi), alpha);
        G[k] = std::fma(0.5 * (prev + cur), h, G[k - 1]);
        prev = cur;
    }
    const double total = G[GRID];
    std::vector<double> out(count);
    size_t k = 0;
    const double denom = count > 1 ? static_cast<double>(count - 1) : 1.0;
    for (size_t t = 0; t < count; ++t) {
        const double target = total > 0.0 ? (static_cast<double>(t) / denom) * total : 0.0;
        while (k < GRID - 1 && G[k + 1] < target) ++k;
        const double g0 = G[k], g1 = G[k + 1];
        const double frac = g1 > g0 ? (target - g0) / (g1 - g0) : 0.0;
        out[t] = std::fma(static_cast<double>(k) + frac, h, j0);
    }
    return out;
}

// The reference's per-step diagnostic CSV (sqa_annealer.cpp:786-798).
struct DebugCsv {
    std::ofstream f;
    DebugCsv(const char *path, const char *tag) {
        if (!path || !path[0]) return;
        f.open(path);
        if (f) {
            f << "phase,step_index,j_perp,chi_B,delta_j,eps_tilde,alpha,floor_hit,chi_B_online\n";
            f.setf(std::ios::scientific);
            f.precision(8);
        } else {
            std::cerr << tag << " WARNING: could not open debug CSV '" << path << "'.\n";
        }
    }
    explicit operator bool() const { return f.is_open() && f.good(); }
};

// chi_B = <B^2> - <B>^2 over a step's samples (no floor); fewer than two samples: the per-bond
// estimate n P (1 - (B/nP)^2) of replica 0's B (sqa_annealer.cpp:729-757, parallel tempering
// :597-615). The single-replica guard of SQAAnnealer raises it to that estimate.
double chi_of(const BondStats &st, double nM, bool single_replica_guard) {
    double var;
    if (st.count >= 2) {
        const double bm = st.sum / static_cast<double>(st.count);
        const double b2m = st.sum2 / static_cast<double>(st.count);
        var = std::fma(-bm, bm, b2m);
    } else {
        const double b_avg = st.b0 / nM;
        var = nM * std::fma(-b_avg, b_avg, 1.0);
    }
    if (single_replica_guard && st.count > 0) {
        const double b_avg = st.b0 / nM;
        var = std::max(var, nM * std::fma(-b_avg, b_avg, 1.0));
    }
    return var;
}

void add_bonds(BondStats &st, const std::vector<int64_t> &b) {
    for (int64_t v : b) {
        const double x = static_cast<double>(v);
        st.sum += x;
        st.sum2 += x * x;
        ++st.count;
    }
}

void write_schedule_scalars(const qa_opt_schedule &o, double eps, double j0, double j1, double jf) {
    if (o.calibrated_eps_tilde) *o.calibrated_eps_tilde = eps;
    if (o.j_perp_start) *o.j_perp_start = j0;
    if (o.resolved_j_perp_end) *o.resolved_j_perp_end = j1;
    if (o.final_j_perp) *o.final_j_perp = jf;
}

}  // namespace

void sqa_optimal(qa_ctx *ctx, const qa_problem *p, const qa_sqa_opt_params *prm, qa_sqa_opt_result *res) {
    const int64_t n = p->n, P = prm->trotter_slices, R = prm->replicas, steps = prm->num_steps;
    const int64_t spb = prm->sweeps_per_step, wl = prm->worldline_sweeps, cl = prm->cluster_sweeps;
    const double alpha = prm->alpha, j0 = prm->j_perp_start, j1 = prm->j_perp_end;
    const bool calibrate = prm->eps_tilde <= 0.0;
    const double nM = static_cast<double>(n * P);
    double floor_chi = 1e-4 * nM;
    double eps = prm->eps_tilde;
    const bool fp32 = prm->mode == QA_MODE_FP32;
    std::unique_ptr<SqaFp32Session> f;
    std::unique_ptr<SqaValSession> v;
    if (fp32)
        f = std::make_unique<SqaFp32Session>(ctx, p, R, P, prm->seed);
    else
        v = std::make_unique<SqaValSession>(ctx, p, R, P, prm->init_spins, prm->replica_seeds);

    // k steps at (bs[q], jp[q]); record: step tail into trace[q]
    auto sweep_steps = [&](int64_t k, const double *bs, const double *jp, int64_t nspb, int64_t ncl, int64_t nwl,
                           bool record, BondStats *st, double *trace) {
        if (f) {
            f->run(k, bs, jp, 0, nullptr, nspb, ncl, nwl, record ? 0 : 2, st, trace);
            return;
        }
        std::vector<int64_t> b;
        for (int64_t q = 0; q < k; ++q) {
            const double join = 1.0 - std::exp(-2.0 * jp[q]);
            BondStats s{};
            for (int64_t w = 0; w < nspb; ++w) {
                v->phase(SqaValSession::kSlice, bs[q], jp[q], join);
                v->bonds(b);
                add_bonds(s, b);
            }
            for (int64_t w = 0; w < ncl; ++w) {
                v->phase(SqaValSession::kCluster, bs[q], jp[q], join);
                v->bonds(b);
                add_bonds(s, b);
            }
            for (int64_t w = 0; w < nwl; ++w) v->phase(SqaValSession::kWorldline, bs[q], jp[q], join);
            if (nspb + ncl == 0) v->bonds(b);
            s.b0 = static_cast<double>(b[0]);
            st[q] = s;
            if (record) trace[q] = v->tail();
        }
    };

    // two phases: beta ramp at j_perp_start, then the adaptive schedule (sqa_annealer.cpp:777-784)
    const double bs_cold = prm->beta / static_cast<double>(P);
    int64_t n1 = 0;
    if (prm->beta_ramp_fraction > 0.0) {
        n1 = static_cast<int64_t>(prm->beta_ramp_fraction * static_cast<double>(steps));
        if (n1 >= steps) n1 = steps - 1;
    }
    const int64_t n2 = steps - n1;
    DebugCsv dbg(prm->debug_csv_path, "[qanneal SQA run_optimal]");
    double *jtr = res->sched.j_perp_trace, *ctr = res->sched.chi_B_trace;

    if (n1 > 0) {  // phase 1 (sqa_annealer.cpp:801-813)
        std::vector<double> bs(n1), jp(n1, j0), beta_t(n1);
        for (int64_t k = 0; k < n1; ++k) {
            const double frac = n1 > 1 ? static_cast<double>(k) / static_cast<double>(n1 - 1) : 1.0;
            beta_t[k] = std::fma(prm->beta - prm->beta_ramp_start, frac, prm->beta_ramp_start);
            bs[k] = beta_t[k] / static_cast<double>(P);
        }
        std::vector<BondStats> st(n1);
        sweep_steps(n1, bs.data(), jp.data(), spb, cl, wl, true, st.data(), res->energy_trace);
        for (int64_t k = 0; k < n1; ++k) {
            const double chi = std::max(chi_of(st[k], nM, R == 1), 0.0);
            jtr[k] = j0;
            ctr[k] = chi;
            if (dbg) dbg.f << "beta_ramp," << k << ',' << j0 << ',' << chi << ",0,0," << alpha << ",0," << beta_t[k] << '\n';
        }
    }

    ChiProfile prof;
    if (calibrate) {  // calibration pilot on a saved copy of the state (sqa_annealer.cpp:820-876)
        const int64_t M = std::max<int64_t>(prm->calib_probes, 2);
        const int64_t cs = std::max<int64_t>(prm->calib_sweeps, 1);
        const int64_t pcl = cl > 0 ? std::max<int64_t>(cl, 1) : 0;
        if (f) {
            if (n1 == 0)  // a launch without sweeps draws the random start state
                f->run(1, &bs_cold, &j0, 0, nullptr, 0, 0, 0, 2, nullptr, nullptr);
            f->save();
        } else {
            v->save();
        }
        const double dJ = j1 - j0;
        const double dJp = dJ / static_cast<double>(M);
        std::vector<double> Jm(M), bs(M, bs_cold);
        for (int64_t m = 0; m < M; ++m) Jm[m] = std::fma(static_cast<double>(m) + 0.5, dJp, j0);
        std::vector<BondStats> st(M);
        sweep_steps(M, bs.data(), Jm.data(), cs, pcl, 0, false, st.data(), nullptr);
        prof.J = Jm;
        prof.chi.assign(M, 0.0);
        double integral = 0.0;
        for (int64_t m = 0; m < M; ++m) {
            double chi = chi_of(st[m], nM, R == 1);
            if (!(chi > 0.0)) chi = 0.0;
            prof.chi[m] = chi;
            integral = std::fma(std::pow(std::max(chi, 1e-12), alpha), dJp, integral);
            if (dbg) dbg.f << "calib," << (m - M) << ',' << Jm[m] << ',' << chi << ",0,0," << alpha << ",0," << chi << '\n';
        }
        if (f) f->restore();
        else v->restore();
        eps = integral / static_cast<double>(n2);
        if (!(eps > 0.0) || !std::isfinite(eps)) {
            std::cerr << "[qanneal SQA run_optimal] WARNING: calibration produced non-positive eps_tilde (integral="
                      << integral << "); falling back to linear ramp.\n";
            eps = std::max(dJ, 0.0) / static_cast<double>(steps);
            if (!(eps > 0.0)) eps = 1e-9;
        }
        const double max_chi = *std::max_element(prof.chi.begin(), prof.chi.end());
        floor_chi = std::max(1e-4 * max_chi, 1e-12);
    }

    // phase 2 (sqa_annealer.cpp:925-949)
    double j = j0;
    std::vector<double> traj;
    if (calibrate && !prof.J.empty() && j1 > j0)
        traj = inverse_cdf_trajectory(prof, j0, j1, floor_chi, alpha, static_cast<size_t>(n2));
    auto book = [&](int64_t k, double jv, const BondStats &st) {  // one phase-2 step's schedule entry
        const double chi_online = chi_of(st, nM, R == 1);
        const double chi_prof = prof.at(jv);
        const double raw = chi_prof >= 0.0 ? chi_prof : chi_online;
        const bool floor_hit = raw < floor_chi;
        const double drive = std::max(raw, floor_chi);
        const double dj = eps * std::pow(drive, -alpha);
        jtr[n1 + k] = jv;
        ctr[n1 + k] = drive;
        if (dbg)
            dbg.f << "run," << (n1 + k) << ',' << jv << ',' << drive << ',' << dj << ',' << eps << ',' << alpha << ','
                  << (floor_hit ? 1 : 0) << ',' << std::max(chi_online, 0.0) << '\n';
        return dj;
    };
    if (!traj.empty()) {
        std::vector<double> bs(n2, bs_cold);
        std::vector<BondStats> st(n2);
        sweep_steps(n2, bs.data(), traj.data(), spb, cl, wl, true, st.data(), res->energy_trace + n1);
        for (int64_t k = 0; k < n2; ++k) book(k, traj[k], st[k]);
        j = traj[n2 - 1];
    } else {
        for (int64_t k = 0; k < n2; ++k) {
            BondStats st;
            sweep_steps(1, &bs_cold, &j, spb, cl, wl, true, &st, res->energy_trace + n1 + k);
            const double dj = book(k, j, st);
            j = std::min(j + dj, j1);
        }
    }
    write_schedule_scalars(res->sched, eps, j0, j1, j);
    if (f) {
        f->best(res->best_state, res->best_energy);
        if (res->kernel_ms) *res->kernel_ms = f->ms;
    } else {
        std::memcpy(res->best_state, v->best.data(), n);
        *res->best_energy = v->best_e;
        if (res->kernel_ms) *res->kernel_ms = v->ms;
    }
}

void sqapt_optimal(qa_ctx *ctx, const qa_problem *p, const qa_sqapt_opt_params *prm, qa_sqapt_opt_result *res) {
    const int64_t n = p->n, P = prm->trotter_slices, R = prm->rungs, steps = prm->num_steps;
    const int64_t spb = prm->sweeps_per_step, wl = prm->worldline_sweeps, cl = prm->cluster_sweeps;
    const double alpha = prm->alpha;
    const bool calibrate = prm->eps_tilde <= 0.0;
    const bool fp32 = prm->mode == QA_MODE_FP32;
    const double nM = static_cast<double>(n * P);
    double eps = prm->eps_tilde;
    const char *tag = "[qanneal run_optimal]";

    // J_perp range (sqa_parallel_tempering.cpp:375-404)
    const double j0 = trotter_coupling(prm->betas[0], prm->gammas[0], P);
    double j = j0, j1 = prm->j_perp_end;
    if (j1 <= 0.0) {
        j1 = trotter_coupling(prm->betas[R - 1], prm->gammas[R - 1], P);
        std::cerr << tag << " WARNING: j_perp_end not specified; defaulted to coldest-replica trotter_coupling = " << j1
                  << " (j_perp_start = " << j0 << "). For benchmark runs pass an explicit j_perp_end (e.g. 5*j_rms).\n";
    }
    if (j1 <= j) {
        std::cerr << tag << " WARNING: resolved j_perp_end (" << j1 << ") <= j_perp_start (" << j0
                  << "). The adaptive schedule has no range to traverse and will not anneal. Installing finite "
                     "ceiling j_perp_start + 1e6 so the run terminates by num_steps. Check your beta/gamma ladder "
                     "or pass an explicit j_perp_end > j_perp_start.\n";
        j1 = j + 1e6;
    }
    if (!fp32) {
        QA_REQUIRE(prm->init_spins && prm->rung_seeds, "validation mode needs init_spins and rung_seeds");
        QA_REQUIRE(!calibrate || prm->scratch_seeds, "validation mode with calibration needs scratch_seeds");
        QA_REQUIRE(prm->swap_uniforms || prm->swap_uniforms_len == 0, "validation mode needs swap_uniforms");
    }
    std::vector<double> bs(R);
    for (int64_t r = 0; r < R; ++r) bs[r] = prm->betas[r] / static_cast<double>(P);

    // Step engine over the ladder: E[r*P + t] slice energies by rung after each step.
    struct Ladder {
        std::unique_ptr<SqaFp32Session> f;
        std::unique_ptr<PtValSession> v;
        std::vector<int> slot, por;  // FP32: rung -> buffer, buffer -> rung
    };
    auto sweep_step = [&](Ladder &L, double jp, int64_t nspb, int64_t ncl, int64_t nwl, BondStats &st,
                          std::vector<double> *E) {
        st = BondStats{};
        if (L.f) {
            std::vector<double> jpv(R, jp);
            L.f->run(1, bs.data(), jpv.data(), R, L.por.data(), nspb, ncl, nwl, 1, &st, nullptr);
            // B of rung 0 from its buffer's chain bond sums
            double b = 0.0;
            for (int64_t t = 0; t < P; ++t) b += static_cast<double>(L.f->chain_b[L.slot[0] * P + t]);
            st.b0 = b;
            if (E) {
                E->resize(R * P);
                for (int64_t r = 0; r < R; ++r)
                    for (int64_t t = 0; t < P; ++t) (*E)[r * P + t] = L.f->chain_e[L.slot[r] * P + t];
            }
            return;
        }
        std::vector<double> jpv(R, jp), join(R, 1.0 - std::exp(-2.0 * jp));
        std::vector<int64_t> b;
        for (int64_t w = 0; w < nspb; ++w) {
            L.v->phase(SqaValSession::kSlice, bs.data(), jpv.data(), join.data());
            L.v->bonds(b);
            add_bonds(st, b);
        }
        for (int64_t w = 0; w < ncl; ++w) {
            L.v->phase(SqaValSession::kCluster, bs.data(), jpv.data(), join.data());
            L.v->bonds(b);
            add_bonds(st, b);
        }
        for (int64_t w = 0; w < nwl; ++w) L.v->phase(SqaValSession::kWorldline, bs.data(), jpv.data(), join.data());
        if (nspb + ncl == 0) L.v->bonds(b);
        st.b0 = static_cast<double>(b[0]);
        if (E) L.v->energies(*E);
    };

    Ladder main;
    std::vector<double> E;
    if (fp32) {
        main.f = std::make_unique<SqaFp32Session>(ctx, p, R, P, prm->seed);
        main.slot.resize(R);
        main.por.resize(R);
        for (int64_t r = 0; r < R; ++r) main.slot[r] = main.por[r] = static_cast<int>(r);
        BondStats st;
        sweep_step(main, j0, 0, 0, 0, st, &E);  // draws the start states and their slice energies
    } else {
        main.v = std::make_unique<PtValSession>(ctx, p, R, P, prm->init_spins, false, prm->rung_seeds);
        main.v->energies(E);
    }

    // projected energy: first strict minimum over the rung's slices (:156-170)
    auto projected = [&](int64_t r, int64_t *bt) {
        double m = std::numeric_limits<double>::infinity();
        int64_t b = 0;
        for (int64_t t = 0; t < P; ++t)
            if (E[r * P + t] < m) {
                m = E[r * P + t];
                b = t;
            }
        *bt = b;
        return m;
    };
    auto slice_state = [&](int64_t r, int64_t t, int8_t *out) {
        if (main.f) {
            QA_HIP(hipMemcpyAsync(out, main.f->device_spins() + (static_cast<int64_t>(main.slot[r]) * P + t) * n, n,
                                  hipMemcpyDeviceToHost, ctx->stream));
            QA_HIP(hipStreamSynchronize(ctx->stream));
        } else {
            main.v->slice_state(r, t, out);
        }
    };
    double best_e = std::numeric_limits<double>::infinity();
    std::vector<int8_t> best(n, 1);
    std::vector<double> en(R);
    auto take_best = [&]() {
        for (int64_t r = 0; r < R; ++r) {
            int64_t bt;
            en[r] = projected(r, &bt);
            if (en[r] < best_e) {
                best_e = en[r];
                slice_state(r, bt, best.data());
            }
        }
    };
    take_best();

    double floor_chi = 1e-4 * nM;
    ChiProfile prof;
    DebugCsv dbg(prm->debug_csv_path, tag);
    if (calibrate) {  // pilot on scratch copies with their own engines (:652-711)
        const int64_t M = std::max<int64_t>(prm->calib_probes, 2);
        const int64_t cs = std::max<int64_t>(prm->calib_sweeps, 1);
        const int64_t pcl = cl > 0 ? std::max<int64_t>(cl, 1) : 0;
        Ladder scratch;
        if (fp32) {
            scratch.f = std::make_unique<SqaFp32Session>(ctx, p, R, P, prm->seed ^ 0x9E3779B97F4A7C15ull);
            scratch.f->copy_state_from(*main.f);
            scratch.slot = main.slot;
            scratch.por = main.por;
        } else {
            // the pilot sweeps copies of the states with its own engines but shuffles the rungs'
            // visit orders in place, which the run then continues from (:454-470, :672)
            scratch.v = std::make_unique<PtValSession>(ctx, p, R, P, main.v->device_spins(), true, prm->scratch_seeds);
            scratch.v->adopt_orders(*main.v);
        }
        const double dJ = j1 - j0;
        const double dJp = dJ / static_cast<double>(M);
        prof.J.resize(M);
        prof.chi.assign(M, 0.0);
        double integral = 0.0;
        for (int64_t m = 0; m < M; ++m) {
            const double Jm = std::fma(static_cast<double>(m) + 0.5, dJp, j0);
            BondStats st;
            sweep_step(scratch, Jm, cs, pcl, 0, st, nullptr);
            double chi = chi_of(st, nM, false);
            if (!(chi > 0.0)) chi = 0.0;
            prof.J[m] = Jm;
            prof.chi[m] = chi;
            integral = std::fma(std::pow(std::max(chi, 1e-12), alpha), dJp, integral);
            if (dbg) dbg.f << "calib," << (m - M) << ',' << Jm << ',' << chi << ",0,0," << alpha << ",0," << chi << '\n';
        }
        if (scratch.v) main.v->adopt_orders(*scratch.v);
        eps = integral / static_cast<double>(steps);
        if (!(eps > 0.0) || !std::isfinite(eps)) {
            std::cerr << tag << " WARNING: calibration produced non-positive eps_tilde (integral=" << integral
                      << "); falling back to linear ramp.\n";
            eps = std::max(dJ, 0.0) / static_cast<double>(steps);
            if (!(eps > 0.0)) eps = 1e-9;
        }
        const double max_chi = *std::max_element(prof.chi.begin(), prof.chi.end());
        floor_chi = std::max(1e-4 * max_chi, 1e-12);
    }

    std::vector<double> traj;
    if (calibrate && !prof.J.empty() && j1 > j0)
        traj = inverse_cdf_trajectory(prof, j0, j1, floor_chi, alpha, static_cast<size_t>(steps));
    std::mt19937_64 swap_eng(prm->seed);  // FP32 swaps
    std::uniform_real_distribution<double> uniform(0.0, 1.0);
    int64_t ucur = 0;
    for (int64_t step = 0; step < steps; ++step) {
        if (!traj.empty()) j = traj[step];
        BondStats st;
        sweep_step(main, j, spb, cl, wl, st, &E);
        take_best();
        const double chi_online = chi_of(st, nM, false);
        const double chi_prof = prof.at(j);
        const double raw = chi_prof >= 0.0 ? chi_prof : chi_online;
        const bool floor_hit = raw < floor_chi;
        const double drive = std::max(raw, floor_chi);
        const double dj = eps * std::pow(drive, -alpha);
        res->sched.j_perp_trace[step] = j;
        res->sched.chi_B_trace[step] = drive;
        if (dbg)
            dbg.f << "run," << step << ',' << j << ',' << drive << ',' << dj << ',' << eps << ',' << alpha << ','
                  << (floor_hit ? 1 : 0) << ',' << std::max(chi_online, 0.0) << '\n';
        if (traj.empty()) j = std::min(j + dj, j1);
        // swaps: the shared J_perp cancels, leaving (beta_r/P - beta_r+1/P)(C_r+1 - C_r) (:812-828)
        double acc = 0.0, att = 0.0;
        if ((step + 1) % prm->swap_interval == 0) {
            for (int64_t r = 0; r + 1 < R; ++r) {
                double c0 = 0.0, c1 = 0.0;
                for (int64_t t = 0; t < P; ++t) c0 += E[r * P + t];
                for (int64_t t = 0; t < P; ++t) c1 += E[(r + 1) * P + t];
                const double delta = (bs[r] - bs[r + 1]) * (c1 - c0);
                att += 1.0;
                bool ok = delta <= 0.0;
                if (!ok) {
                    double u;
                    if (fp32) {
                        u = uniform(swap_eng);
                    } else {
                        if (ucur >= prm->swap_uniforms_len) throw Error{QA_EINVAL, "swap uniform stream exhausted"};
                        u = prm->swap_uniforms[ucur++];
                    }
                    ok = u < std::exp(-delta);
                }
                if (ok) {
                    if (fp32) {
                        std::swap(main.slot[r], main.slot[r + 1]);
                        main.por[main.slot[r]] = static_cast<int>(r);
                        main.por[main.slot[r + 1]] = static_cast<int>(r + 1);
                    } else {
                        main.v->swap(r);
                    }
                    std::swap_ranges(E.begin() + r * P, E.begin() + (r + 1) * P, E.begin() + (r + 1) * P);
                    std::swap(en[r], en[r + 1]);
                    acc += 1.0;
                }
            }
        }
        double avg = 0.0;
        for (double e : en) avg += e;
        res->average_energy_trace[step] = avg / static_cast<double>(R);
        res->swap_acceptance_trace[step] = att > 0.0 ? acc / att : 0.0;
    }
    write_schedule_scalars(res->sched, eps, j0, j1, j);
    for (int64_t r = 0; r < R; ++r) {
        int64_t bt;
        res->final_energies[r] = projected(r, &bt);
        slice_state(r, bt, res->final_states + r * n);
    }
    std::memcpy(res->best_state, best.data(), n);
    *res->best_energy = best_e;
    if (fp32) {  // FP32: report the exact FP64 energies of the returned states
        DevBuf<int8_t> st((R + 1) * n);
        DevBuf<double> ex(R + 1);
        st.upload(res->final_states, R * n, ctx->stream);
        QA_HIP(hipMemcpyAsync(st.ptr + R * n, best.data(), n, hipMemcpyHostToDevice, ctx->stream));
        device_energies(p, st.ptr, R + 1, ex.ptr, ctx->stream);
        std::vector<double> hx(R + 1);
        ex.download(hx.data(), R + 1, ctx->stream);
        QA_HIP(hipStreamSynchronize(ctx->stream));
        for (int64_t r = 0; r < R; ++r) res->final_energies[r] = hx[r];
        *res->best_energy = hx[R];
    }
    if (res->swap_uniforms_used) *res->swap_uniforms_used = ucur;
    if (res->kernel_ms) *res->kernel_ms = main.f ? main.f->ms : main.v->ms;
}

}  // namespace qa
