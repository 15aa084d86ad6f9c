// sqa_chi (SQAChiAnnealer::run_chi, sqa_chi_annealer.cpp:146-438): pilot chi_B scan, schedule
// inversion, production anneal.
//
// FP32 mode: all scan points run in one launch of the FP32 sweep kernel (each point = R
// independent replicas on a fresh random worldline with its own J_perp), then the production
// anneal runs in a second launch. FP64 validation mode replays the reference: every (replica,
// slice) cell owns an engine seeded from the caller's draws, parity phases sweep the cells one
// wave per cell, and the pilot statistics and schedule are accumulated on the host in the
// reference's order and rounding.
//
// The schedule construction is host FP64 arithmetic (a 4096-point trapezoid integral and an
// inverse-CDF lookup); it is shared by both modes and restates :283-378 with the reference
// build's contraction sites as explicit fma (everything else uncontracted).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "qa_internal.hpp"
#include "qa_rng.hpp"

#pragma clang fp contract(off)

namespace qa {

void sqa_fp32_pilot(qa_ctx *ctx, const qa_problem *p, int64_t P, int64_t R, int64_t points,
                    const std::vector<double> &bs, const std::vector<double> &jp, int64_t burn, int64_t sweeps,
                    uint64_t seed, std::vector<double> &m2, std::vector<double> &mabs, float *ms);
void sqa_fp32_chi_production(qa_ctx *ctx, const qa_problem *p, int64_t P, int64_t R, int64_t steps,
                             const double *betas, const double *gammas, int64_t sweeps, uint64_t seed,
                             int8_t *best_state, double *best_energy, double *trace, float *ms);

namespace {

struct ChiGrid {
    double A0, s_lo, s_hi;
    double g2s(double g) const { return A0 / (A0 + g); }
    double s2g(double s) const {
        s = std::min(std::max(s, 1e-10), 1.0 - 1e-10);
        return A0 * (1.0 - s) / s;
    }
};

ChiGrid make_grid(const qa_chi_params *prm) {
    ChiGrid g;
    g.A0 = prm->driver_A0 > 0.0 ? prm->driver_A0 : prm->gamma_start;
    g.s_lo = g.g2s(prm->gamma_start);
    g.s_hi = g.g2s(prm->gamma_end);
    return g;
}

// scan grid s_q uniform in [s_lo, s_hi] (:403-407)
void scan_grid(const qa_chi_params *prm, const ChiGrid &g, int64_t P, std::vector<double> &s,
               std::vector<double> &gam, std::vector<double> &jp) {
    const int64_t SP = prm->scan_points;
    s.resize(SP);
    gam.resize(SP);
    jp.resize(SP);
    for (int64_t q = 0; q < SP; ++q) {
        s[q] = g.s_lo + (g.s_hi - g.s_lo) * static_cast<double>(q) / static_cast<double>(SP - 1);
        gam[q] = g.s2g(s[q]);
        jp[q] = trotter_coupling(prm->beta, gam[q], P);
    }
}

// chi_B peak, floor, trapezoid cumulative weight, phase-1 beta ramp and inverse-CDF gamma(t)
void build_schedule(const qa_chi_params *prm, const ChiGrid &g, int64_t P, const std::vector<double> &ss,
                    const std::vector<double> &sg, const std::vector<double> &chi, qa_chi_result *res) {
    const int64_t SP = prm->scan_points, NS = prm->num_steps;
    for (int64_t q = 0; q < SP; ++q) {
        res->scan_s[q] = ss[q];
        res->scan_gamma[q] = sg[q];
        res->scan_chi_B[q] = chi[q];
    }
    const int64_t peak = std::max_element(chi.begin(), chi.end()) - chi.begin();
    res->scalars[0] = ss[peak];
    res->scalars[1] = sg[peak];
    res->scalars[2] = trotter_coupling(prm->beta, sg[peak], P);
    const double mx = chi[peak];
    const bool degenerate = !(mx > 0.0) || !std::isfinite(mx);
    const double floor = degenerate ? 1.0 : std::max(prm->chi_floor_fraction * mx, 1e-12);
    res->scalars[3] = floor;
    res->scalars[4] = g.A0;
    auto weight = [&](double s) {
        double c;
        if (s <= ss[0]) c = chi[0];
        else if (s >= ss[SP - 1]) c = chi[SP - 1];
        else {
            int64_t hi = 1;
            while (hi < SP && ss[hi] < s) ++hi;
            const int64_t lo = hi - 1;
            const double t = (s - ss[lo]) / (ss[hi] - ss[lo]);
            c = std::fma(t, chi[hi] - chi[lo], chi[lo]);
        }
        return std::max(c, floor);
    };
    int64_t n1 = 0;
    if (prm->beta_ramp_fraction > 0.0) {
        n1 = static_cast<int64_t>(prm->beta_ramp_fraction * static_cast<double>(NS));
        if (n1 > NS - 2) n1 = NS - 2;
    }
    const int64_t n2 = NS - n1;
    const int64_t GRID = 4096;
    const double hs = (g.s_hi - g.s_lo) / static_cast<double>(GRID);
    std::vector<double> W(GRID + 1, 0.0);
    double pw = weight(g.s_lo);
    for (int64_t k = 1; k <= GRID; ++k) {
        const double cw = weight(std::fma(static_cast<double>(k), hs, g.s_lo));
        W[k] = std::fma(0.5 * (pw + cw), hs, W[k - 1]);
        pw = cw;
    }
    const double Wt = W[GRID];
    for (int64_t k = 0; k < NS; ++k) {
        res->beta_schedule[k] = prm->beta;
        res->gamma_schedule[k] = prm->gamma_start;
        res->s_schedule[k] = g.s_lo;
    }
    for (int64_t k = 0; k < n1; ++k) {
        const double fr = n1 > 1 ? static_cast<double>(k) / static_cast<double>(n1 - 1) : 1.0;
        res->beta_schedule[k] = std::fma(prm->beta - prm->beta_ramp_start, fr, prm->beta_ramp_start);
    }
    int64_t ki = 0;
    for (int64_t t = 0; t < n2; ++t) {
        const double target = Wt > 0.0 ? (static_cast<double>(t) / static_cast<double>(n2 - 1)) * Wt : 0.0;
        while (ki < GRID && W[ki + 1] < target) ++ki;
        const double w0 = W[ki], w1 = W[ki + 1];
        const double fr = w1 > w0 ? (target - w0) / (w1 - w0) : 0.0;
        const double s_t = std::fma(static_cast<double>(ki) + fr, hs, g.s_lo);
        res->s_schedule[n1 + t] = s_t;
        res->gamma_schedule[n1 + t] = std::min(std::max(g.s2g(s_t), prm->gamma_end), prm->gamma_start);
    }
}

// ---- validation: one parity phase, one wave per cell -------------------------------------
struct ChiPhaseArgs {
    ProblemView pv;
    int P;
    double bs, jp;
    int8_t *spins;          // [R][P][n]
    double *fields;         // [R][P][n]
    const int32_t *cells;   // cell ids (r*P + t) of this phase
    const uint64_t *stream; // [cells][n]
    int64_t *used;          // [cells]
};

__device__ void update_dense_fp64(const ProblemView &pv, double *f, int flip, int8_t old, int lane) {
    const double ds = -2.0 * static_cast<double>(old);
    const double *row = pv.J64 + static_cast<int64_t>(flip) * pv.n;
    for (int j = lane; j < pv.n; j += 64) {
        double v = f[j] + row[j] * ds;
        if (j == flip) v -= row[flip] * ds;
        f[j] = v;
    }
}

__global__ void __launch_bounds__(64) chi_phase_kernel(ChiPhaseArgs a) {
    const int lane = threadIdx.x;
    const int cell = a.cells[blockIdx.x];
    const int n = a.pv.n, P = a.P;
    const int r = cell / P, t = cell - r * P;
    int8_t *cur = a.spins + static_cast<int64_t>(cell) * n;
    double *f = a.fields + static_cast<int64_t>(cell) * n;
    const int8_t *prv = a.spins + static_cast<int64_t>(r * P + (t == 0 ? P - 1 : t - 1)) * n;
    const int8_t *nxt = a.spins + static_cast<int64_t>(r * P + (t + 1) % P) * n;
    const uint64_t *str = a.stream + static_cast<int64_t>(blockIdx.x) * n;
    int64_t used = 0;
    for (int i = 0; i < n; ++i) {
        const int8_t s = cur[i];
        const int nn = static_cast<int>(prv[i]) + static_cast<int>(nxt[i]);
        const double dcl = -2.0 * static_cast<double>(s) * f[i];
        const double dtr = 2.0 * a.jp * static_cast<double>(s) * static_cast<double>(nn);
        const double de = __fma_rn(a.bs, dcl, dtr);  // sqa_chi_annealer.cpp:115
        bool acc = de <= 0.0;
        if (!acc) acc = canonical_from_u64(str[used++]) < exp(-de);
        if (acc) {
            __syncthreads();
            if (lane == 0) cur[i] = static_cast<int8_t>(-s);
            __syncthreads();
            if (a.pv.kind == QA_KIND_DENSE) {
                update_dense_fp64(a.pv, f, i, s, lane);
            } else if (lane == 0) {
                if (a.pv.kind == QA_KIND_SPARSE) {
                    const double ds = -2.0 * static_cast<double>(s);
                    for (int q = a.pv.adj_ptr[i]; q < a.pv.adj_ptr[i + 1]; ++q) f[a.pv.adj_idx[q]] += a.pv.adj_val64[q] * ds;
                } else {
                    for (int q = a.pv.inc_ptr[i]; q < a.pv.inc_ptr[i + 1]; ++q) {
                        const int tt = a.pv.inc_idx[q];
                        double prod = a.pv.t_coef64[tt];
                        for (int k = a.pv.t_off[tt]; k < a.pv.t_off[tt + 1]; ++k) prod *= static_cast<double>(cur[a.pv.t_vars[k]]);
                        for (int k = a.pv.t_off[tt]; k < a.pv.t_off[tt + 1]; ++k) {
                            const int v = a.pv.t_vars[k];
                            if (v != i) f[v] += 2.0 * prod * static_cast<double>(cur[v]);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    if (lane == 0) a.used[blockIdx.x] = used;
}

// per-replica spin sums (int64), one block per replica
__global__ void replica_sums(const int8_t *spins, int P, int n, long long *out) {
    const int8_t *S = spins + static_cast<int64_t>(blockIdx.x) * P * n;
    long long acc = 0;
    for (int e = threadIdx.x; e < P * n; e += blockDim.x) acc += S[e];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ long long part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long s = 0;
        for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) s += part[w];
        out[blockIdx.x] = s;
    }
}

}  // namespace

void chi_validation(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, qa_chi_result *res) {
    QA_REQUIRE(prm->init_spins && prm->cell_seeds, "validation mode needs init_spins and cell_seeds");
    QA_REQUIRE(!prm->chi_override && !prm->surrogate, "a supplied or predicted chi_B profile is an FP32-mode path");
    const int64_t n = p->n, R = prm->replicas, P = prm->trotter_slices, SP = prm->scan_points, NS = prm->num_steps;
    const int64_t cells = R * P;
    for (int64_t k = 0; k < (SP + 1) * cells * n; ++k)
        QA_REQUIRE(prm->init_spins[k] == 1 || prm->init_spins[k] == -1, "State spins must be -1 or +1.");
    hipStream_t s = ctx->stream;
    const ChiGrid g = make_grid(prm);
    std::vector<double> ss, sg, sjp;
    scan_grid(prm, g, P, ss, sg, sjp);
    const double bs = prm->beta / static_cast<double>(P);
    DevBuf<int8_t> spins(cells * n);
    DevBuf<double> fields(cells * n), energies(cells);
    DevBuf<int32_t> dcells(cells);
    DevBuf<uint64_t> stream(cells * n);
    DevBuf<int64_t> used(cells);
    DevBuf<long long> rsum(R);
    std::vector<std::mt19937_64> eng(cells);
    std::vector<uint64_t> hstream(cells * n);
    std::vector<int64_t> hused(cells);
    std::vector<long long> hsum(R);
    std::vector<double> mv(R), hen(cells);
    // parity phases in the reference's single-thread order: even slices (for odd P without the
    // last slice), then the last slice when P is odd, then odd slices; cells r-major within a phase
    std::vector<std::vector<int32_t>> phases;
    {
        std::vector<int32_t> ev, last, od;
        for (int64_t r = 0; r < R; ++r)
            for (int64_t t = 0; t < P; ++t) {
                const int32_t c = static_cast<int32_t>(r * P + t);
                if (t % 2 == 1) od.push_back(c);
                else if ((P % 2 == 1) && t == P - 1 && P > 1) last.push_back(c);
                else ev.push_back(c);
            }
        phases.push_back(ev);
        if (!last.empty()) phases.push_back(last);
        if (!od.empty()) phases.push_back(od);
    }
    float total_ms = 0.0f;
    auto fresh = [&](int64_t block) {
        spins.upload(prm->init_spins + block * cells * n, cells * n, s);
        fields_fp64_launch(p, spins.ptr, fields.ptr, cells, s);
        for (int64_t c = 0; c < cells; ++c) eng[c].seed(prm->cell_seeds[block * cells + c]);
    };
    auto sweep = [&](double bsv, double jpv) {
        for (const auto &ph : phases) {
            const int64_t m = static_cast<int64_t>(ph.size());
            for (int64_t k = 0; k < m; ++k) {
                std::mt19937_64 copy = eng[ph[k]];
                for (int64_t i = 0; i < n; ++i) hstream[k * n + i] = copy();
            }
            dcells.upload(ph.data(), m, s);
            stream.upload(hstream.data(), m * n, s);
            ChiPhaseArgs a{view_of(p), static_cast<int>(P), bsv, jpv, spins.ptr, fields.ptr, dcells.ptr, stream.ptr, used.ptr};
            QA_HIP(hipEventRecord(ctx->ev0, s));
            hipLaunchKernelGGL(chi_phase_kernel, dim3(static_cast<unsigned>(m)), dim3(64), 0, s, a);
            QA_HIP(hipGetLastError());
            QA_HIP(hipEventRecord(ctx->ev1, s));
            used.download(hused.data(), m, s);
            QA_HIP(hipStreamSynchronize(s));
            float ms = 0;
            QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            total_ms += ms;
            for (int64_t k = 0; k < m; ++k) eng[ph[k]].discard(static_cast<unsigned long long>(hused[k]));
        }
        hipLaunchKernelGGL(replica_sums, dim3(static_cast<unsigned>(R)), dim3(256), 0, s, spins.ptr,
                           static_cast<int>(P), static_cast<int>(n), rsum.ptr);
        QA_HIP(hipGetLastError());
        rsum.download(hsum.data(), R, s);
        QA_HIP(hipStreamSynchronize(s));
        const double inv = 1.0 / static_cast<double>(n * P);
        for (int64_t r = 0; r < R; ++r) mv[r] = static_cast<double>(hsum[r]) * inv;
    };
    std::vector<double> chi(SP, 0.0);
    for (int64_t q = 0; q < SP; ++q) {  // pilot scan (:232-281)
        fresh(q);
        double m2 = 0.0, ma = 0.0;
        int64_t cnt = 0;
        for (int64_t k = 0; k < prm->scan_burn + prm->scan_sweeps; ++k) {
            sweep(bs, sjp[q]);
            if (k >= prm->scan_burn)
                for (int64_t j = 0; j < R; ++j) {
                    // the reference build sums four replicas' rounded squares at a time and
                    // contracts only the R % 4 tail (see oracle/orc_restated.cpp)
                    m2 = (j < (R / 4) * 4) ? m2 + mv[j] * mv[j] : std::fma(mv[j], mv[j], m2);
                    ma += std::fabs(mv[j]);
                    ++cnt;
                }
        }
        if (cnt >= 2) {
            const double m2m = m2 / static_cast<double>(cnt), mam = ma / static_cast<double>(cnt);
            chi[q] = std::max(static_cast<double>(n * P) * std::fma(-mam, mam, m2m), 0.0);
        }
    }
    build_schedule(prm, g, P, ss, sg, chi, res);
    fresh(SP);  // production on a fresh worldline (:394-405)
    double best_e = std::numeric_limits<double>::infinity();
    std::vector<int8_t> best(n, 1);
    for (int64_t k = 0; k < NS; ++k) {
        const double bsk = res->beta_schedule[k] / static_cast<double>(P);
        const double jpk = trotter_coupling(res->beta_schedule[k], res->gamma_schedule[k], P);
        for (int64_t q = 0; q < prm->sweeps_per_step; ++q) sweep(bsk, jpk);
        device_energies(p, spins.ptr, cells, energies.ptr, s);
        energies.download(hen.data(), cells, s);
        QA_HIP(hipStreamSynchronize(s));
        double avg = 0.0;
        int64_t arg = -1;
        for (int64_t c = 0; c < cells; ++c) {
            avg += hen[c];
            if (hen[c] < best_e) { best_e = hen[c]; arg = c; }
        }
        if (arg >= 0) {
            QA_HIP(hipMemcpyAsync(best.data(), spins.ptr + arg * n, n, hipMemcpyDeviceToHost, s));
            QA_HIP(hipStreamSynchronize(s));
        }
        res->energy_trace[k] = avg / static_cast<double>(cells);
    }
    std::memcpy(res->best_state, best.data(), n);
    *res->best_energy = best_e;
    if (res->kernel_ms) *res->kernel_ms = total_ms;
}

namespace {

// chi_B of each scan point from the FP32 pilot (:232-281): chi = max(n P (<m^2> - <|m|>^2), 0)
void pilot_chi(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, const std::vector<double> &sjp,
               std::vector<double> &chi, float *ms) {
    const int64_t R = prm->replicas, P = prm->trotter_slices, SP = prm->scan_points;
    std::vector<double> bsv(SP, prm->beta / static_cast<double>(P)), m2, ma;
    sqa_fp32_pilot(ctx, p, P, R, SP, bsv, sjp, prm->scan_burn, prm->scan_sweeps, prm->seed, m2, ma, ms);
    const double cnt = static_cast<double>(prm->scan_sweeps * R);
    chi.assign(SP, 0.0);
    for (int64_t q = 0; q < SP; ++q) {
        double s2 = 0.0, sa = 0.0;
        for (int64_t r = 0; r < R; ++r) {
            s2 += m2[q * R + r];
            sa += ma[q * R + r];
        }
        if (cnt >= 2.0) {
            const double m2m = s2 / cnt, mam = sa / cnt;
            chi[q] = std::max(static_cast<double>(p->n * P) * (m2m - mam * mam), 0.0);
        }
    }
}

// Surrogate profile: ensemble mean of the normalized chi_B prediction. Returns false (pilot
// fallback) when the features leave the training range by more than a quarter of its width or
// the largest per-point standard deviation across members exceeds the caller's bound.
bool surrogate_chi(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, std::vector<double> &chi,
                   double *spread) {
    qa_mlp *m = const_cast<qa_mlp *>(prm->surrogate);
    const int64_t SP = prm->scan_points;
    QA_REQUIRE(m->ctx == ctx, "the surrogate belongs to another context");
    QA_REQUIRE(m->trained, "the surrogate has not been trained or loaded");
    QA_REQUIRE(m->n_in == QA_CHI_NFEATURES && m->n_out == SP,
               "the surrogate's input/output sizes do not match the features and scan_points");
    double f[QA_CHI_NFEATURES];
    chi_features(p, prm->beta, prm->trotter_slices, prm->driver_A0, prm->gamma_start, prm->gamma_end, f);
    *spread = std::numeric_limits<double>::infinity();
    for (int k = 0; k < QA_CHI_NFEATURES; ++k) {
        const double v = (f[k] - m->x_mean[k]) / m->x_scale[k];
        const double w = 0.25 * (m->x_hi[k] - m->x_lo[k]);
        if (!(v >= m->x_lo[k] - w && v <= m->x_hi[k] + w)) return false;
    }
    std::vector<float> y(static_cast<size_t>(m->members) * SP);
    mlp_forward(m, 1, f, y.data());
    chi.assign(SP, 0.0);
    double worst = 0.0;
    for (int64_t q = 0; q < SP; ++q) {
        double mu = 0.0, sq = 0.0;
        for (int z = 0; z < m->members; ++z) mu += y[z * SP + q];
        mu /= m->members;
        for (int z = 0; z < m->members; ++z) sq += (y[z * SP + q] - mu) * (y[z * SP + q] - mu);
        chi[q] = mu;
        worst = std::max(worst, std::sqrt(sq / m->members));
    }
    *spread = worst;
    return worst <= prm->surrogate_max_spread;
}

}  // namespace

void chi_features(const qa_problem *p, double beta, int64_t P, double A0, double gamma_start, double gamma_end,
                  double *f) {
    const double F = std::max(p->st_field_scale, 1e-12);
    const double A = A0 > 0.0 ? A0 : gamma_start;
    f[0] = std::log2(static_cast<double>(p->n));
    f[1] = std::log2(1.0 + p->st_mean_degree);
    f[2] = std::log(F);
    f[3] = p->st_field_het;
    f[4] = std::log(std::max(p->st_kurtosis, 1e-12));
    f[5] = p->st_neg_frac;
    f[6] = p->st_h_ratio;
    f[7] = p->st_order;
    f[8] = std::log(beta * F);
    f[9] = std::log(static_cast<double>(P));
    f[10] = std::log(A / F);
    f[11] = std::log(gamma_end / gamma_start);
}

void chi_pilot_profile(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, double *scan_s, double *scan_gamma,
                       double *scan_chi, float *ms) {
    const ChiGrid g = make_grid(prm);
    std::vector<double> ss, sg, sjp, chi;
    scan_grid(prm, g, prm->trotter_slices, ss, sg, sjp);
    pilot_chi(ctx, p, prm, sjp, chi, ms);
    for (int64_t q = 0; q < prm->scan_points; ++q) {
        scan_s[q] = ss[q];
        scan_gamma[q] = sg[q];
        scan_chi[q] = chi[q];
    }
}

void chi_fp32_schedule(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, qa_chi_result *res) {
    const int64_t P = prm->trotter_slices, SP = prm->scan_points;
    const ChiGrid g = make_grid(prm);
    std::vector<double> ss, sg, sjp;
    scan_grid(prm, g, P, ss, sg, sjp);
    std::vector<double> chi(SP, 0.0);
    float ms_pilot = 0.0f;
    int32_t used = 0;
    double spread = std::numeric_limits<double>::quiet_NaN();
    if (prm->chi_override) {
        for (int64_t q = 0; q < SP; ++q) chi[q] = prm->chi_override[q];
    } else if (prm->surrogate && surrogate_chi(ctx, p, prm, chi, &spread)) {
        used = 1;
    } else {
        pilot_chi(ctx, p, prm, sjp, chi, &ms_pilot);
    }
    build_schedule(prm, g, P, ss, sg, chi, res);
    if (res->kernel_ms) *res->kernel_ms = ms_pilot;
    if (res->surrogate_used) *res->surrogate_used = used;
    if (res->surrogate_spread) *res->surrogate_spread = spread;
}

void chi_fp32(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, qa_chi_result *res) {
    double ms_pilot = 0.0;
    qa_chi_result head = *res;
    head.kernel_ms = &ms_pilot;
    chi_fp32_schedule(ctx, p, prm, &head);
    float ms_prod = 0.0f;
    sqa_fp32_chi_production(ctx, p, prm->trotter_slices, prm->replicas, prm->num_steps, res->beta_schedule,
                            res->gamma_schedule, prm->sweeps_per_step, prm->seed ^ 0x9E3779B97F4A7C15ull,
                            res->best_state, res->best_energy, res->energy_trace, &ms_prod);
    if (res->kernel_ms) *res->kernel_ms = static_cast<float>(ms_pilot) + ms_prod;
}

}  // namespace qa
