// ORACLE — TEST INFRASTRUCTURE ONLY (never linked into the product).
//
// A CPU restatement of the reference annealing engines, written from the reference's documented
// behaviour: FP64 throughout, libstdc++ <random> (std::mt19937_64, uniform_real_distribution,
// uniform_int_distribution, std::shuffle) with the reference's draw order, and OpenMP over
// replicas / rungs / parity cells exactly where the reference forks. Each function cites the
// reference file:line it follows.
//
// Floating-point contraction. The reference is built with GCC -O3 -march=native (its
// CMakeLists.txt:69-83), and GCC contracts a*b+c into FMA on x86-64 with FMA. This file is built
// with -ffp-contract=off and calls std::fma explicitly at every site where the reference build
// contracts a product whose result is not exact; products of an FP64 value with +-1 or +-2 are
// exact, so contraction there cannot change a result and is left as plain arithmetic. The site
// list was read off the reference objects' disassembly and is pinned by tests/golden fixtures:
//   sqa_annealer.cpp:200,275       de = fma(beta/P, dE_classical, dE_trotter)
//   sqa_parallel_tempering.cpp:153 action = fma(beta/P, sum E_t, -(J_perp * B))
//   sqa_parallel_tempering.cpp:215,268
//   sqa_chi_annealer.cpp:115,253 (R % 4 tail only),265,327,344,346,359,374.
//
// Third-party arithmetic (not part of /root/reference): libstdc++ of GCC 11.4.0 for <random>, and
// glibc libm for exp/log/tanh. Both are called, not re-implemented.
#include "orc_abi.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
using Rng = std::mt19937_64;

// ---------------------------------------------------------------------------------------------
// Problem model (dense_ising.cpp, sparse_ising.cpp, higher_order_ising.cpp)
// ---------------------------------------------------------------------------------------------
struct Model {
    int kind = ORC_DENSE;
    std::size_t n = 0;
    double c = 0.0;
    std::vector<double> h;
    std::vector<double> J;  // dense row-major n*n (diagonal kept; the dynamics never read it)
    struct Edge { std::size_t i, j; double v; };
    std::vector<Edge> edges;                                   // sparse, caller order
    std::vector<std::vector<std::pair<std::size_t, double>>> adj;  // sparse adjacency
    std::vector<double> tcoef;                                 // hubo consolidated terms
    std::vector<std::size_t> toff;
    std::vector<std::uint32_t> tvars;
    std::vector<std::vector<std::uint32_t>> inc;               // hubo incidence

    static void check_spins(const int8_t *s, std::size_t n) {
        // state.hpp:43-50 validate_spins
        for (std::size_t i = 0; i < n; ++i)
            if (s[i] != 1 && s[i] != -1) throw std::invalid_argument("State spins must be -1 or +1.");
    }

    // dense_ising.cpp:29-44 / sparse_ising.cpp:43-56 / higher_order_ising.cpp:92-111
    double energy(const int8_t *s) const {
        check_spins(s, n);
        double e = c;
        for (std::size_t i = 0; i < n; ++i) e += h[i] * static_cast<double>(s[i]);
        if (kind == ORC_DENSE) {
            for (std::size_t i = 0; i < n; ++i) {
                const double si = static_cast<double>(s[i]);
                const double *row = &J[i * n];
                for (std::size_t j = i + 1; j < n; ++j) e += row[j] * si * static_cast<double>(s[j]);
            }
        } else if (kind == ORC_SPARSE) {
            for (const Edge &ed : edges)
                e += ed.v * static_cast<double>(s[ed.i]) * static_cast<double>(s[ed.j]);
        } else {
            for (std::size_t t = 0; t + 1 < toff.size(); ++t) {
                double prod = tcoef[t];
                for (std::size_t k = toff[t]; k < toff[t + 1]; ++k) prod *= static_cast<double>(s[tvars[k]]);
                e += prod;
            }
        }
        return e;
    }

    // dense_ising.cpp:46-62 / sparse_ising.cpp:58-71 / higher_order_ising.cpp:113-136
    double delta(const int8_t *s, std::size_t f) const {
        if (f >= n) throw std::invalid_argument("Flip index out of range.");
        double local = h[f];
        if (kind == ORC_DENSE) {
            const double *row = &J[f * n];
            for (std::size_t j = 0; j < n; ++j)
                if (j != f) local += row[j] * static_cast<double>(s[j]);
        } else if (kind == ORC_SPARSE) {
            for (const auto &nb : adj[f]) local += nb.second * static_cast<double>(s[nb.first]);
        } else {
            for (std::uint32_t t : inc[f]) {
                double prod = tcoef[t];
                for (std::size_t k = toff[t]; k < toff[t + 1]; ++k)
                    if (tvars[k] != f) prod *= static_cast<double>(s[tvars[k]]);
                local += prod;
            }
        }
        return -2.0 * static_cast<double>(s[f]) * local;
    }

    // dense_ising.cpp:64-83 / sparse_ising.cpp:73-85 / higher_order_ising.cpp:138-160
    void fields(const int8_t *s, double *out) const {
        if (kind == ORC_DENSE) {
            for (std::size_t i = 0; i < n; ++i) {
                const double *row = &J[i * n];
                double f = h[i];
                for (std::size_t j = 0; j < n; ++j) f += row[j] * static_cast<double>(s[j]);
                out[i] = f - row[i] * static_cast<double>(s[i]);
            }
        } else if (kind == ORC_SPARSE) {
            for (std::size_t i = 0; i < n; ++i) out[i] = h[i];
            for (const Edge &ed : edges) {
                out[ed.i] += ed.v * static_cast<double>(s[ed.j]);
                out[ed.j] += ed.v * static_cast<double>(s[ed.i]);
            }
        } else {
            for (std::size_t i = 0; i < n; ++i) out[i] = h[i];
            for (std::size_t t = 0; t + 1 < toff.size(); ++t) {
                double prod = tcoef[t];
                for (std::size_t k = toff[t]; k < toff[t + 1]; ++k) prod *= static_cast<double>(s[tvars[k]]);
                for (std::size_t k = toff[t]; k < toff[t + 1]; ++k)
                    out[tvars[k]] += prod * static_cast<double>(s[tvars[k]]);
            }
        }
    }

    // dense_ising.cpp:88-106 / sparse_ising.cpp:90-102 / higher_order_ising.cpp:167-189
    void update(double *f, const int8_t *ns, std::size_t flip, int8_t old) const {
        if (kind == ORC_DENSE) {
            const double ds = -2.0 * static_cast<double>(old);
            const double *row = &J[flip * n];
            for (std::size_t j = 0; j < n; ++j) f[j] += row[j] * ds;
            // The reference then subtracts the self term again: f[flip] = (f + x) - x with
            // x = 2*J_ff*(-old). That is not an exact restore when J_ff != 0 (QUBO-derived J has a
            // nonzero diagonal), so the same two roundings are reproduced here.
            f[flip] -= row[flip] * ds;
        } else if (kind == ORC_SPARSE) {
            const double ds = -2.0 * static_cast<double>(old);
            for (const auto &nb : adj[flip]) f[nb.first] += nb.second * ds;
        } else {
            for (std::uint32_t t : inc[flip]) {
                double prod = tcoef[t];
                for (std::size_t k = toff[t]; k < toff[t + 1]; ++k) prod *= static_cast<double>(ns[tvars[k]]);
                for (std::size_t k = toff[t]; k < toff[t + 1]; ++k)
                    if (tvars[k] != flip) f[tvars[k]] += 2.0 * prod * static_cast<double>(ns[tvars[k]]);
            }
        }
    }
};

Model build(const orc_problem *p) {
    Model m;
    m.kind = p->kind;
    if (p->n <= 0) throw std::invalid_argument("problem size must be > 0.");
    m.n = static_cast<std::size_t>(p->n);
    m.c = p->c;
    const std::size_t n = m.n;
    if (p->kind == ORC_DENSE) {
        m.h.assign(p->h, p->h + n);
        m.J.assign(p->J, p->J + n * n);
    } else if (p->kind == ORC_SPARSE) {
        m.h.assign(p->h, p->h + n);
        m.adj.assign(n, {});
        for (int64_t e = 0; e < p->nedges; ++e) {
            // sparse_ising.cpp:17-41: range / self-edge checks, then adjacency in edge order
            const std::size_t i = static_cast<std::size_t>(p->ei[e]), j = static_cast<std::size_t>(p->ej[e]);
            if (i >= n || j >= n) throw std::invalid_argument("SparseIsing edge index out of range.");
            if (i == j) throw std::invalid_argument("SparseIsing self-edge not allowed.");
            m.edges.push_back({i, j, p->ev[e]});
        }
        for (const auto &ed : m.edges) {
            m.adj[ed.i].push_back({ed.j, ed.v});
            m.adj[ed.j].push_back({ed.i, ed.v});
        }
    } else {
        // higher_order_ising.cpp:10-90: parity reduction, fold degree 0/1, merge equal supports in
        // the order of an ordered map keyed by the sorted variable list.
        m.h.assign(n, 0.0);
        std::map<std::vector<std::uint32_t>, double> merged;
        for (int64_t t = 0; t < p->nterms; ++t) {
            const double coeff = p->term_coeff[t];
            if (coeff == 0.0) continue;
            std::vector<std::size_t> v(p->term_vars + p->term_off[t], p->term_vars + p->term_off[t + 1]);
            for (std::size_t x : v)
                if (x >= n) throw std::invalid_argument("HigherOrderIsing term variable index out of range.");
            std::sort(v.begin(), v.end());
            std::vector<std::uint32_t> odd;
            for (std::size_t a = 0; a < v.size();) {
                std::size_t b = a;
                while (b < v.size() && v[b] == v[a]) ++b;
                if ((b - a) % 2 == 1) odd.push_back(static_cast<std::uint32_t>(v[a]));
                a = b;
            }
            if (odd.empty()) m.c += coeff;
            else if (odd.size() == 1) m.h[odd[0]] += coeff;
            else merged[odd] += coeff;
        }
        m.toff.push_back(0);
        m.inc.assign(n, {});
        for (const auto &kv : merged) {
            if (kv.second == 0.0) continue;
            const std::uint32_t id = static_cast<std::uint32_t>(m.tcoef.size());
            m.tcoef.push_back(kv.second);
            for (std::uint32_t x : kv.first) {
                m.tvars.push_back(x);
                m.inc[x].push_back(id);
            }
            m.toff.push_back(m.tvars.size());
        }
    }
    return m;
}

// sqa_annealer.cpp:66-70 (identical copies at sqa_parallel_tempering.cpp:75-79, sqa_chi_annealer.cpp:49-53)
double trotter_coupling(double beta, double gamma, std::size_t slices) {
    const double x = std::max(beta * gamma / static_cast<double>(slices), 1e-12);
    return 0.5 * std::log(1.0 / std::tanh(x));
}

// sqa_state.hpp:64-72 / state.hpp:22-30: one uniform_int<int>(0,1) draw per spin, flat order.
void random_spins(std::vector<int8_t> &s, Rng &rng) {
    std::uniform_int_distribution<int> bit(0, 1);
    for (auto &v : s) v = bit(rng) ? 1 : -1;
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

// One replica's imaginary-time worldline sweeps at fixed (beta_scale, J_perp), shared by the
// SQA engine (cached fields) — sqa_annealer.cpp:176-315.
struct SqaReplica {
    const Model *m;
    std::size_t n, P;
    int8_t *spins;     // [P][n]
    double *fields;    // [P][n]
    std::vector<std::size_t> *order;
    std::vector<char> *in_cluster;
    std::vector<std::size_t> *stack;

    int8_t *sl(std::size_t t) const { return spins + t * n; }
    double *fl(std::size_t t) const { return fields + t * n; }

    void slice_sweep(Rng &rng, double bs, double jp) {
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        std::shuffle(order->begin(), order->end(), rng);
        for (std::size_t t = 0; t < P; ++t) {
            int8_t *cur = sl(t);
            double *f = fl(t);
            const int8_t *prv = sl(t == 0 ? P - 1 : t - 1);
            const int8_t *nxt = sl((t + 1) % P);
            for (std::size_t k = 0; k < n; ++k) {
                const std::size_t i = (*order)[k];
                const int8_t s = cur[i];
                const double dcl = -2.0 * static_cast<double>(s) * f[i];
                const double dtr = 2.0 * jp * static_cast<double>(s) * static_cast<double>(prv[i] + nxt[i]);
                const double de = std::fma(bs, dcl, dtr);  // sqa_annealer.cpp:200
                if (de <= 0.0 || uni(rng) < std::exp(-de)) {
                    cur[i] = static_cast<int8_t>(-s);
                    m->update(f, cur, i, s);
                }
            }
        }
    }

    void cluster_sweep(Rng &rng, double bs, double jp, double join) {
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        std::uniform_int_distribution<std::size_t> pick(0, P - 1);
        std::vector<char> &inc = *in_cluster;
        std::vector<std::size_t> &stk = *stack;
        std::shuffle(order->begin(), order->end(), rng);
        for (std::size_t k = 0; k < n; ++k) {
            const std::size_t i = (*order)[k];
            const std::size_t seed = pick(rng);
            std::fill(inc.begin(), inc.end(), 0);
            stk.clear();
            stk.push_back(seed);
            inc[seed] = 1;
            const int8_t s0 = sl(seed)[i];
            while (!stk.empty()) {  // LIFO growth along the ring, prev before next (:228-246)
                const std::size_t t = stk.back();
                stk.pop_back();
                const std::size_t pv = t == 0 ? P - 1 : t - 1, nx = (t + 1) % P;
                if (!inc[pv] && sl(pv)[i] == s0 && uni(rng) < join) { inc[pv] = 1; stk.push_back(pv); }
                if (!inc[nx] && sl(nx)[i] == s0 && uni(rng) < join) { inc[nx] = 1; stk.push_back(nx); }
            }
            double dcl = 0.0, dtm = 0.0;
            for (std::size_t t = 0; t < P; ++t) {
                if (!inc[t]) continue;
                dcl += -2.0 * static_cast<double>(s0) * fl(t)[i];
                const std::size_t pv = t == 0 ? P - 1 : t - 1, nx = (t + 1) % P;
                if (!inc[pv]) dtm += 2.0 * jp * static_cast<double>(s0) * static_cast<double>(sl(pv)[i]);
                if (!inc[nx]) dtm += 2.0 * jp * static_cast<double>(s0) * static_cast<double>(sl(nx)[i]);
            }
            const double de = std::fma(bs, dcl, dtm);  // sqa_annealer.cpp:275
            if (de <= 0.0 || uni(rng) < std::exp(-de)) {
                for (std::size_t t = 0; t < P; ++t) {
                    if (!inc[t]) continue;
                    sl(t)[i] = static_cast<int8_t>(-s0);
                    m->update(fl(t), sl(t), i, s0);
                }
            }
        }
    }

    void worldline_sweep(Rng &rng, double bs) {
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        std::shuffle(order->begin(), order->end(), rng);
        for (std::size_t k = 0; k < n; ++k) {
            const std::size_t i = (*order)[k];
            double dcl = 0.0;
            for (std::size_t t = 0; t < P; ++t) dcl += -2.0 * static_cast<double>(sl(t)[i]) * fl(t)[i];
            const double de = bs * dcl;
            if (de <= 0.0 || uni(rng) < std::exp(-de)) {
                for (std::size_t t = 0; t < P; ++t) {
                    const int8_t old = sl(t)[i];
                    sl(t)[i] = static_cast<int8_t>(-old);
                    m->update(fl(t), sl(t), i, old);
                }
            }
        }
    }
};

// One rung's sweeps in SQA parallel tempering: no field cache, the classical dE of every attempt
// is the O(n) delta_energy — sqa_parallel_tempering.cpp:208-296 (run) and :468-588 (run_optimal,
// same arithmetic with one shared J_perp).
struct PtReplica {
    const Model *m;
    std::size_t n, P;
    int8_t *spins;  // [P][n]
    std::vector<std::size_t> *order;
    std::vector<char> *in_cluster;
    std::vector<std::size_t> *stack;
    bool shuffle = true;  // false: sites in index order (SQAAnnealer's tracing path, :317-453)

    int8_t *sl(std::size_t t) const { return spins + t * n; }

    void slice_sweep(Rng &g, double bs, double jp) {
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        if (shuffle) std::shuffle(order->begin(), order->end(), g);
        for (std::size_t t = 0; t < P; ++t) {
            int8_t *cur = sl(t);
            const int8_t *prv = sl(t == 0 ? P - 1 : t - 1), *nxt = sl((t + 1) % P);
            for (std::size_t q = 0; q < n; ++q) {
                const std::size_t i = (*order)[q];
                const double dcl = m->delta(cur, i);
                const double dtr = 2.0 * jp * static_cast<double>(cur[i]) * static_cast<double>(prv[i] + nxt[i]);
                const double dt = std::fma(bs, dcl, dtr);  // :215
                if (dt <= 0.0 || uni(g) < std::exp(-dt)) cur[i] = static_cast<int8_t>(-cur[i]);
            }
        }
    }

    void cluster_sweep(Rng &g, double bs, double jp, double join) {
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        std::uniform_int_distribution<std::size_t> pick(0, P - 1);
        std::vector<char> &inc = *in_cluster;
        std::vector<std::size_t> &sk = *stack;
        if (shuffle) std::shuffle(order->begin(), order->end(), g);
        for (std::size_t q = 0; q < n; ++q) {
            const std::size_t i = (*order)[q];
            const std::size_t sd = pick(g);
            std::fill(inc.begin(), inc.end(), 0);
            sk.clear();
            sk.push_back(sd);
            inc[sd] = 1;
            const int8_t s0 = sl(sd)[i];
            while (!sk.empty()) {
                const std::size_t t = sk.back();
                sk.pop_back();
                const std::size_t pv = t == 0 ? P - 1 : t - 1, nx = (t + 1) % P;
                if (!inc[pv] && sl(pv)[i] == s0 && uni(g) < join) { inc[pv] = 1; sk.push_back(pv); }
                if (!inc[nx] && sl(nx)[i] == s0 && uni(g) < join) { inc[nx] = 1; sk.push_back(nx); }
            }
            double dcl = 0.0, dtm = 0.0;
            for (std::size_t t = 0; t < P; ++t) {
                if (!inc[t]) continue;
                dcl += m->delta(sl(t), i);
                const std::size_t pv = t == 0 ? P - 1 : t - 1, nx = (t + 1) % P;
                const double st_ = static_cast<double>(sl(t)[i]);
                if (!inc[pv]) dtm += 2.0 * jp * st_ * static_cast<double>(sl(pv)[i]);
                if (!inc[nx]) dtm += 2.0 * jp * st_ * static_cast<double>(sl(nx)[i]);
            }
            const double dt = std::fma(bs, dcl, dtm);  // :268
            if (dt <= 0.0 || uni(g) < std::exp(-dt))
                for (std::size_t t = 0; t < P; ++t)
                    if (inc[t]) sl(t)[i] = static_cast<int8_t>(-sl(t)[i]);
        }
    }

    void worldline_sweep(Rng &g, double bs) {
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        if (shuffle) std::shuffle(order->begin(), order->end(), g);
        for (std::size_t q = 0; q < n; ++q) {
            const std::size_t i = (*order)[q];
            double dcl = 0.0;
            for (std::size_t t = 0; t < P; ++t) dcl += m->delta(sl(t), i);
            const double dt = bs * dcl;
            if (dt <= 0.0 || uni(g) < std::exp(-dt))
                for (std::size_t t = 0; t < P; ++t) sl(t)[i] = static_cast<int8_t>(-sl(t)[i]);
        }
    }
};

// B = sum_t sum_i s_t,i s_t+1,i of one replica (sqa_annealer.cpp:483-497): an exact integer.
double bond_sum(const int8_t *s, std::size_t P, std::size_t n) {
    double sum = 0.0;
    for (std::size_t t = 0; t < P; ++t) {
        const int8_t *a = s + t * n, *b = s + ((t + 1) % P) * n;
        for (std::size_t i = 0; i < n; ++i) sum += static_cast<double>(a[i]) * static_cast<double>(b[i]);
    }
    return sum;
}

// Var(B) over every replica's samples, no floor (sqa_annealer.cpp:729-757; parallel tempering
// :597-615): b2m - bm*bm contracts to one fma in the reference build. fallback_b: B used when
// fewer than two samples exist.
double var_b(const std::vector<std::vector<double>> &samples, double nM, double fallback_b) {
    double sum = 0.0, sum2 = 0.0;
    std::size_t total = 0;
    for (const auto &v : samples)
        for (double b : v) {
            sum += b;
            sum2 += b * b;
            ++total;
        }
    if (total >= 2) {
        const double bm = sum / static_cast<double>(total);
        const double b2m = sum2 / static_cast<double>(total);
        return std::fma(-bm, bm, b2m);
    }
    const double b_avg = fallback_b / nM;
    return nM * std::fma(-b_avg, b_avg, 1.0);
}

// chi_B(J) between calibration probes, linear and clamped; -1 with no probes (:880-891).
double profile_at(const std::vector<double> &J, const std::vector<double> &chi, double x) {
    const std::size_t M = J.size();
    if (M == 0) return -1.0;
    if (x <= J.front()) return chi.front();
    if (x >= J.back()) return chi.back();
    std::size_t hi = 1;
    while (hi < M && J[hi] < x) ++hi;
    const std::size_t lo = hi - 1;
    const double t = (x - J[lo]) / (J[hi] - J[lo]);
    return std::fma(t, chi[hi] - chi[lo], chi[lo]);
}

// Inverse-CDF trajectory over a 4096-cell trapezoid grid of max(chi, floor)^alpha (:895-921).
std::vector<double> trajectory(const std::vector<double> &J, const std::vector<double> &chi, double j0, double j1,
                               double floor_chi, double alpha, std::size_t count) {
    const std::size_t GRID = 4096;
    const double h = (j1 - j0) / static_cast<double>(GRID);
    std::vector<double> G(GRID + 1, 0.0);
    double prev = std::pow(std::max(profile_at(J, chi, j0), floor_chi), alpha);
    for (std::size_t k = 1; k <= GRID; ++k) {
        const double Jk = std::fma(static_cast<double>(k), h, j0);
        const double cur = std::pow(std::max(profile_at(J, chi, Jk), floor_chi), alpha);
        G[k] = std::fma(0.5 * (prev + cur), h, G[k - 1]);
        prev = cur;
    }
    const double total = G[GRID];
    std::vector<double> out(count);
    std::size_t k = 0;
    const double denom = count > 1 ? static_cast<double>(count - 1) : 1.0;
    for (std::size_t t = 0; t < count; ++t) {
        const double target = total > 0.0 ? (static_cast<double>(t) / denom) * total : 0.0;
        while (k < GRID && G[k + 1] < target) ++k;
        const double g0 = G[k], g1 = G[k + 1];
        const double frac = g1 > g0 ? (target - g0) / (g1 - g0) : 0.0;
        out[t] = std::fma(static_cast<double>(k) + frac, h, j0);
    }
    return out;
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
        const Model m = build(p);
        for (int64_t k = 0; k < count; ++k) out[k] = m.energy(spins + k * p->n);
    });
}

int orc_delta_energy(const orc_problem *p, const int8_t *spins, int64_t flip, double *out) {
    return guard([&] { *out = build(p).delta(spins, static_cast<std::size_t>(flip)); });
}

int orc_local_fields(const orc_problem *p, const int8_t *spins, double *fields) {
    return guard([&] { build(p).fields(spins, fields); });
}

int orc_update_fields(const orc_problem *p, double *fields, const int8_t *new_spins, int64_t flip,
                      int8_t old_spin) {
    return guard([&] { build(p).update(fields, new_spins, static_cast<std::size_t>(flip), old_spin); });
}

// qubo.cpp:33-69
int orc_qubo_to_ising(int64_t n64, const double *Q, double *h, double *J, double *c) {
    return guard([&] {
        if (n64 <= 0) throw std::invalid_argument("QUBO size must be > 0.");
        const std::size_t n = static_cast<std::size_t>(n64);
        std::vector<double> W(n * n);
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j < n; ++j) W[i * n + j] = 0.5 * (Q[i * n + j] + Q[j * n + i]);
        for (std::size_t i = 0; i < n; ++i) {
            double rs = 0.0;
            for (std::size_t j = 0; j < n; ++j) rs += W[i * n + j];
            h[i] = 0.5 * rs;
        }
        double cc = 0.0;
        for (std::size_t i = 0; i < n; ++i) {
            cc += 0.25 * W[i * n + i];
            for (std::size_t j = 0; j < n; ++j) {
                J[i * n + j] = 0.5 * W[i * n + j];
                cc += 0.25 * W[i * n + j];
            }
        }
        *c = cc;
    });
}

// annealer.cpp:33-75 (set_seed: annealer.cpp:29-31)
int orc_sa_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
               uint64_t seed, int8_t *best_state, double *best_energy, double *trace) {
    return guard([&] {
        if (steps <= 0) throw std::invalid_argument("Schedule must contain at least one beta.");
        if (sweeps_per_beta <= 0) throw std::invalid_argument("sweeps_per_beta must be > 0.");
        const Model m = build(p);
        const std::size_t n = m.n;
        Rng rng(seed);
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        std::vector<int8_t> s(n);
        random_spins(s, rng);
        double e = m.energy(s.data());
        std::vector<int8_t> best = s;
        double best_e = e;
        for (int64_t step = 0; step < steps; ++step) {
            const double beta = betas[step];
            for (int64_t sw = 0; sw < sweeps_per_beta; ++sw) {
                for (std::size_t i = 0; i < n; ++i) {
                    const double d = m.delta(s.data(), i);
                    if (d <= 0.0 || uni(rng) < std::exp(-beta * d)) {
                        s[i] = static_cast<int8_t>(-s[i]);
                        e += d;
                        if (e < best_e) {  // best after every improving accepted flip (:58-61)
                            best_e = e;
                            best = s;
                        }
                    }
                }
            }
            trace[step] = e;
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
    });
}

// replica_annealer.cpp:46-142: all replicas' random states first (:55-59), then one engine per
// replica seeded from the master (:60-63); sites in index order; best after every improving
// accepted flip of each replica; global best and averages at step ends (:113-137).
int orc_replica_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
                    int64_t replicas, uint64_t seed, int8_t *best_states, double *best_energies,
                    double *energy_traces, double *mag_traces, int8_t *global_best_state,
                    double *global_best_energy, double *avg_energy_trace, double *avg_mag_trace) {
    return guard([&] {
        if (sweeps_per_beta <= 0) throw std::invalid_argument("sweeps_per_beta must be > 0.");
        if (replicas <= 0) throw std::invalid_argument("replicas must be > 0.");
        const Model m = build(p);
        const std::size_t n = m.n, R = static_cast<std::size_t>(replicas);
        Rng rng(seed);
        std::vector<std::vector<int8_t>> st(R, std::vector<int8_t>(n));
        std::vector<double> e(R);
        for (std::size_t r = 0; r < R; ++r) {
            random_spins(st[r], rng);
            e[r] = m.energy(st[r].data());
        }
        std::vector<Rng> rr(R);
        for (std::size_t r = 0; r < R; ++r) rr[r].seed(rng());
        std::vector<std::vector<int8_t>> best = st;
        std::vector<double> best_e = e;
        std::vector<int8_t> gbest = st[0];
        double gbest_e = std::numeric_limits<double>::infinity();
        for (std::size_t r = 0; r < R; ++r)
            if (e[r] < gbest_e) { gbest_e = e[r]; gbest = st[r]; }
        auto mag = [&](const std::vector<int8_t> &s) {
            double sum = 0.0;
            for (int8_t v : s) sum += static_cast<double>(v);
            return n ? sum / static_cast<double>(n) : 0.0;
        };
        for (int64_t step = 0; step < steps; ++step) {
            const double beta = betas[step];
            for (std::size_t r = 0; r < R; ++r) {
                std::uniform_real_distribution<double> uni(0.0, 1.0);
                for (int64_t sw = 0; sw < sweeps_per_beta; ++sw)
                    for (std::size_t i = 0; i < n; ++i) {
                        const double d = m.delta(st[r].data(), i);
                        if (d <= 0.0 || uni(rr[r]) < std::exp(-beta * d)) {
                            st[r][i] = static_cast<int8_t>(-st[r][i]);
                            e[r] += d;
                            if (e[r] < best_e[r]) { best_e[r] = e[r]; best[r] = st[r]; }
                        }
                    }
            }
            double ae = 0.0, am = 0.0;
            for (std::size_t r = 0; r < R; ++r) {
                if (best_e[r] < gbest_e) { gbest_e = best_e[r]; gbest = best[r]; }
            }
            for (std::size_t r = 0; r < R; ++r) {
                const double mg = mag(st[r]);
                energy_traces[r * steps + step] = e[r];
                mag_traces[r * steps + step] = mg;
                ae += e[r];
                am += mg;
            }
            avg_energy_trace[step] = ae / static_cast<double>(R);
            avg_mag_trace[step] = am / static_cast<double>(R);
        }
        for (std::size_t r = 0; r < R; ++r) {
            std::memcpy(best_states + r * n, best[r].data(), n);
            best_energies[r] = best_e[r];
        }
        std::memcpy(global_best_state, gbest.data(), n);
        *global_best_energy = gbest_e;
    });
}

// parallel_tempering.cpp:36-122: one master engine for everything -- the random states in replica
// order (:56-59), then every replica's sweeps in replica order and the swap chain, all drawing
// uniforms from the same engine; best checked after each replica's sweeps of a step (:87-90).
int orc_pt_run(const orc_problem *p, int64_t replicas, const double *betas, int64_t sweeps_per_step,
               int64_t steps, int64_t swap_interval, uint64_t seed, int8_t *final_states,
               double *final_energies, int8_t *best_state, double *best_energy, double *avg_trace,
               double *swap_trace) {
    return guard([&] {
        if (sweeps_per_step <= 0) throw std::invalid_argument("sweeps_per_step must be > 0.");
        if (steps <= 0) throw std::invalid_argument("steps must be > 0.");
        if (swap_interval <= 0) throw std::invalid_argument("swap_interval must be > 0.");
        const Model m = build(p);
        const std::size_t n = m.n, R = static_cast<std::size_t>(replicas);
        Rng rng(seed);
        std::vector<std::vector<int8_t>> st(R, std::vector<int8_t>(n));
        std::vector<double> e(R);
        for (std::size_t r = 0; r < R; ++r) {
            random_spins(st[r], rng);
            e[r] = m.energy(st[r].data());
        }
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        std::vector<int8_t> best(n);
        double best_e = std::numeric_limits<double>::infinity();
        for (int64_t step = 0; step < steps; ++step) {
            for (std::size_t r = 0; r < R; ++r) {
                const double beta = betas[r];
                for (int64_t sw = 0; sw < sweeps_per_step; ++sw)
                    for (std::size_t i = 0; i < n; ++i) {
                        const double d = m.delta(st[r].data(), i);
                        if (d <= 0.0 || uni(rng) < std::exp(-beta * d)) {
                            st[r][i] = static_cast<int8_t>(-st[r][i]);
                            e[r] += d;
                        }
                    }
                if (e[r] < best_e) { best_e = e[r]; best = st[r]; }
            }
            double acc = 0.0, att = 0.0;
            if ((step + 1) % swap_interval == 0) {
                for (std::size_t r = 0; r + 1 < R; ++r) {
                    const double delta = (betas[r] - betas[r + 1]) * (e[r + 1] - e[r]);
                    att += 1.0;
                    if (delta <= 0.0 || uni(rng) < std::exp(-delta)) {
                        std::swap(st[r], st[r + 1]);
                        std::swap(e[r], e[r + 1]);
                        acc += 1.0;
                    }
                }
            }
            double avg = 0.0;
            for (double v : e) avg += v;
            avg_trace[step] = avg / static_cast<double>(R);
            swap_trace[step] = att > 0.0 ? acc / att : 0.0;
        }
        for (std::size_t r = 0; r < R; ++r) {
            std::memcpy(final_states + r * n, st[r].data(), n);
            final_energies[r] = e[r];
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
    });
}

// sqa_annealer.cpp:87-477 (the cached-field, observer-free path)
namespace {
// One SQAAnnealer::run on the annealer's engine `rng` (sqa_annealer.cpp:87-477): the engine is
// left where the run leaves it, so a second call continues the stream as a second run() does.
void sqa_run_on(const orc_problem *p, Rng &rng, int64_t steps, const double *betas, const double *gammas,
                int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta, int64_t worldline_sweeps,
                int64_t cluster_sweeps, int64_t continuous_time_slices, int8_t *best_state, double *best_energy,
                double *trace) {
    {
        if (steps <= 0) throw std::invalid_argument("SQA schedule must contain betas.");
        if (trotter_slices <= 0) throw std::invalid_argument("trotter_slices must be > 0.");
        if (replicas <= 0) throw std::invalid_argument("replicas must be > 0.");
        if (sweeps_per_beta == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
            throw std::invalid_argument(
                "At least one of sweeps_per_beta, worldline_sweeps, or cluster_sweeps must be > 0.");
        const Model m = build(p);
        const std::size_t n = m.n, R = static_cast<std::size_t>(replicas);
        const std::size_t P = (continuous_time_slices > trotter_slices)
                                  ? static_cast<std::size_t>(continuous_time_slices)
                                  : static_cast<std::size_t>(trotter_slices);
        std::vector<int8_t> spins(R * P * n);
        random_spins(spins, rng);  // :100, before the replica seeds
        std::vector<Rng> rrng(R);
        for (auto &g : rrng) g.seed(rng());  // :108-111
        std::vector<double> fields(R * P * n);
        for (std::size_t k = 0; k < R * P; ++k) m.fields(&spins[k * n], &fields[k * n]);
        std::vector<std::vector<std::size_t>> order(R, std::vector<std::size_t>(n));
        for (auto &o : order) std::iota(o.begin(), o.end(), std::size_t{0});
        std::vector<std::vector<char>> incl(R, std::vector<char>(P));
        std::vector<std::vector<std::size_t>> stk(R);

        double best_e = std::numeric_limits<double>::infinity();
        std::vector<int8_t> best(n, 1);
        for (int64_t step = 0; step < steps; ++step) {
            const double beta = betas[step], gamma = gammas[step];
            const double jp = trotter_coupling(beta, gamma, P);
            const double bs = beta / static_cast<double>(P);
            const double join = 1.0 - std::exp(-2.0 * jp);
#pragma omp parallel for if (R > 1) schedule(static)
            for (long long rr = 0; rr < static_cast<long long>(R); ++rr) {
                const std::size_t r = static_cast<std::size_t>(rr);
                SqaReplica rep{&m, n, P, &spins[r * P * n], &fields[r * P * n], &order[r], &incl[r], &stk[r]};
                for (int64_t k = 0; k < sweeps_per_beta; ++k) rep.slice_sweep(rrng[r], bs, jp);
                for (int64_t k = 0; k < cluster_sweeps; ++k) rep.cluster_sweep(rrng[r], bs, jp, join);
                for (int64_t k = 0; k < worldline_sweeps; ++k) rep.worldline_sweep(rrng[r], bs);
            }
            // :455-473 serial FP64 energy of every slice; best at step ends only
            double avg = 0.0;
            for (std::size_t k = 0; k < R * P; ++k) {
                const double e = m.energy(&spins[k * n]);
                avg += e;
                if (e < best_e) {
                    best_e = e;
                    std::memcpy(best.data(), &spins[k * n], n);
                }
            }
            trace[step] = avg / static_cast<double>(R * P);
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
    }
}
}  // namespace

int orc_sqa_run(const orc_problem *p, int64_t steps, const double *betas, const double *gammas,
                int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                int64_t worldline_sweeps, int64_t cluster_sweeps, int64_t continuous_time_slices,
                uint64_t seed, int8_t *best_state, double *best_energy, double *trace) {
    return guard([&] {
        Rng rng(seed);
        sqa_run_on(p, rng, steps, betas, gammas, trotter_slices, replicas, sweeps_per_beta, worldline_sweeps,
                   cluster_sweeps, continuous_time_slices, best_state, best_energy, trace);
    });
}

// set_seed(seed), then `runs` consecutive run() calls on the same annealer; outputs of the last.
int orc_sqa_run_repeat(const orc_problem *p, int64_t runs, int64_t steps, const double *betas,
                       const double *gammas, int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                       int64_t worldline_sweeps, int64_t cluster_sweeps, int64_t continuous_time_slices,
                       uint64_t seed, int8_t *best_state, double *best_energy, double *trace) {
    return guard([&] {
        if (runs <= 0) throw std::invalid_argument("runs must be > 0.");
        Rng rng(seed);
        for (int64_t k = 0; k < runs; ++k)
            sqa_run_on(p, rng, steps, betas, gammas, trotter_slices, replicas, sweeps_per_beta, worldline_sweeps,
                       cluster_sweeps, continuous_time_slices, best_state, best_energy, trace);
    });
}

// sqa_parallel_tempering.cpp:81-346
int orc_sqapt_run(const orc_problem *p, int64_t rungs, const double *betas, const double *gammas,
                  int64_t trotter_slices, int64_t sweeps_per_step, int64_t worldline_sweeps,
                  int64_t steps, int64_t swap_interval, int64_t cluster_sweeps,
                  int64_t continuous_time_slices, uint64_t seed, int8_t *final_states,
                  double *final_energies, int8_t *best_state, double *best_energy,
                  double *avg_trace, double *swap_trace) {
    return guard([&] {
        if (rungs < 2) throw std::invalid_argument("SQA parallel tempering requires >=2 beta/gamma pairs.");
        if (trotter_slices <= 0) throw std::invalid_argument("trotter_slices must be > 0.");
        if (steps <= 0) throw std::invalid_argument("steps must be > 0.");
        if (swap_interval <= 0) throw std::invalid_argument("swap_interval must be > 0.");
        if (sweeps_per_step == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
            throw std::invalid_argument(
                "At least one of sweeps_per_step/worldline_sweeps/cluster_sweeps must be > 0.");
        const Model m = build(p);
        const std::size_t n = m.n, R = static_cast<std::size_t>(rungs);
        const std::size_t P = (continuous_time_slices > trotter_slices)
                                  ? static_cast<std::size_t>(continuous_time_slices)
                                  : static_cast<std::size_t>(trotter_slices);
        Rng rng(seed);
        std::vector<std::vector<int8_t>> st(R, std::vector<int8_t>(P * n));
        for (auto &s : st) random_spins(s, rng);  // :103-107
        std::vector<Rng> rrng(R);
        for (auto &g : rrng) g.seed(rng());  // :109-112
        std::vector<std::vector<std::size_t>> order(R, std::vector<std::size_t>(n));
        for (auto &o : order) std::iota(o.begin(), o.end(), std::size_t{0});
        std::vector<std::vector<char>> incl(R, std::vector<char>(P));
        std::vector<std::vector<std::size_t>> stk(R);

        auto proj = [&](const std::vector<int8_t> &s, std::size_t *best_t) {  // :156-170
            double b = std::numeric_limits<double>::infinity();
            std::size_t bt = 0;
            for (std::size_t t = 0; t < P; ++t) {
                const double e = m.energy(&s[t * n]);
                if (e < b) { b = e; bt = t; }
            }
            *best_t = bt;
            return b;
        };
        auto action = [&](const std::vector<int8_t> &s, double beta, double gamma) {  // :128-154
            double cs = 0.0;
            for (std::size_t t = 0; t < P; ++t) cs += m.energy(&s[t * n]);
            double bsum = 0.0;
            for (std::size_t t = 0; t < P; ++t) {
                const int8_t *a = &s[t * n], *b = &s[((t + 1) % P) * n];
                for (std::size_t i = 0; i < n; ++i) bsum += static_cast<double>(a[i]) * static_cast<double>(b[i]);
            }
            const double jp = trotter_coupling(beta, gamma, P);
            return std::fma(beta / static_cast<double>(P), cs, -(jp * bsum));  // :153
        };

        double best_e = std::numeric_limits<double>::infinity();
        std::vector<int8_t> best(n, 1);
        std::vector<double> en(R);
        for (std::size_t r = 0; r < R; ++r) {
            std::size_t bt;
            en[r] = proj(st[r], &bt);
            if (en[r] < best_e) { best_e = en[r]; std::memcpy(best.data(), &st[r][bt * n], n); }
        }
        for (int64_t step = 0; step < steps; ++step) {
#pragma omp parallel for if (R > 1) schedule(static)
            for (long long rl = 0; rl < static_cast<long long>(R); ++rl) {
                const std::size_t r = static_cast<std::size_t>(rl);
                const double bs = betas[r] / static_cast<double>(P);
                const double jp = trotter_coupling(betas[r], gammas[r], P);
                const double join = 1.0 - std::exp(-2.0 * jp);
                PtReplica rep{&m, n, P, st[r].data(), &order[r], &incl[r], &stk[r]};
                for (int64_t k = 0; k < sweeps_per_step; ++k) rep.slice_sweep(rrng[r], bs, jp);      // :208-221
                for (int64_t k = 0; k < cluster_sweeps; ++k) rep.cluster_sweep(rrng[r], bs, jp, join);  // :223-278
                for (int64_t k = 0; k < worldline_sweeps; ++k) rep.worldline_sweep(rrng[r], bs);    // :280-296
            }
            for (std::size_t r = 0; r < R; ++r) {  // :299-306
                std::size_t bt;
                en[r] = proj(st[r], &bt);
                if (en[r] < best_e) { best_e = en[r]; std::memcpy(best.data(), &st[r][bt * n], n); }
            }
            double acc = 0.0, att = 0.0;
            if ((step + 1) % swap_interval == 0) {  // :308-325, master rng
                std::uniform_real_distribution<double> uni(0.0, 1.0);
                for (std::size_t r = 0; r + 1 < R; ++r) {
                    const double eaa = action(st[r], betas[r], gammas[r]);
                    const double ebb = action(st[r + 1], betas[r + 1], gammas[r + 1]);
                    const double eab = action(st[r + 1], betas[r], gammas[r]);
                    const double eba = action(st[r], betas[r + 1], gammas[r + 1]);
                    const double d = (eab + eba) - (eaa + ebb);
                    att += 1.0;
                    if (d <= 0.0 || uni(rng) < std::exp(-d)) {
                        std::swap(st[r], st[r + 1]);
                        std::swap(en[r], en[r + 1]);
                        acc += 1.0;
                    }
                }
            }
            double avg = 0.0;
            for (double e : en) avg += e;
            avg_trace[step] = avg / static_cast<double>(R);
            swap_trace[step] = att > 0.0 ? acc / att : 0.0;
        }
        for (std::size_t r = 0; r < R; ++r) {
            std::size_t bt;
            final_energies[r] = proj(st[r], &bt);
            std::memcpy(final_states + r * n, &st[r][bt * n], n);
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
    });
}

// sqa_chi_annealer.cpp:146-438 (checkerboard sweep :81-140)
int orc_sqa_chi_run(const orc_problem *p, int64_t trotter_slices, int64_t replicas, double beta,
                    double gamma_start, double gamma_end, int64_t num_steps,
                    int64_t sweeps_per_step, int64_t scan_points, int64_t scan_sweeps,
                    int64_t scan_burn, double chi_floor_fraction, double driver_A0,
                    double beta_ramp_fraction, double beta_ramp_start, uint64_t seed,
                    int8_t *best_state, double *best_energy, double *trace, double *beta_sched,
                    double *gamma_sched, double *s_sched, double *scan_s, double *scan_gamma,
                    double *scan_chi, double *scalars) {
    return guard([&] {
        if (trotter_slices <= 0) throw std::invalid_argument("trotter_slices must be > 0.");
        if (replicas <= 0) throw std::invalid_argument("replicas must be > 0.");
        if (num_steps < 2) throw std::invalid_argument("num_steps must be >= 2.");
        if (sweeps_per_step <= 0) throw std::invalid_argument("sweeps_per_step must be > 0.");
        if (!(beta > 0.0)) throw std::invalid_argument("beta must be > 0.");
        if (!(gamma_end > 0.0) || !(gamma_start > gamma_end))
            throw std::invalid_argument("Require gamma_start > gamma_end > 0.");
        if (scan_points < 4) throw std::invalid_argument("scan_points must be >= 4.");
        if (scan_sweeps <= 0) throw std::invalid_argument("scan_sweeps must be > 0.");
        if (chi_floor_fraction < 0.0) throw std::invalid_argument("chi_floor_fraction must be >= 0.");
        const Model m = build(p);
        const std::size_t n = m.n, R = static_cast<std::size_t>(replicas);
        const std::size_t P = static_cast<std::size_t>(trotter_slices);
        const double A0 = driver_A0 > 0.0 ? driver_A0 : gamma_start;
        auto g2s = [A0](double g) { return A0 / (A0 + g); };
        auto s2g = [A0](double s) {
            s = std::min(std::max(s, 1e-10), 1.0 - 1e-10);
            return A0 * (1.0 - s) / s;
        };
        const double s_lo = g2s(gamma_start), s_hi = g2s(gamma_end);
        Rng rng(seed);

        // build_parity_cells (:55-65): (r,t) grouped by t%2, r-major
        std::vector<std::pair<std::size_t, std::size_t>> cells[2];
        for (std::size_t r = 0; r < R; ++r)
            for (std::size_t t = 0; t < P; ++t) cells[t % 2].push_back({r, t});

        std::vector<int8_t> spins(R * P * n);
        std::vector<double> fields(R * P * n);
        std::vector<Rng> crng(R * P);
        std::vector<double> mvals(R);

        auto sweep = [&](double bs, double jp) {  // :81-140
            for (int par = 0; par < 2; ++par) {
                const auto &cl = cells[par];
#pragma omp parallel for if (cl.size() > 1) schedule(static)
                for (long long ci = 0; ci < static_cast<long long>(cl.size()); ++ci) {
                    const std::size_t r = cl[ci].first, t = cl[ci].second;
                    Rng &g = crng[r * P + t];
                    std::uniform_real_distribution<double> uni(0.0, 1.0);
                    int8_t *cur = &spins[(r * P + t) * n];
                    double *f = &fields[(r * P + t) * n];
                    const int8_t *prv = &spins[(r * P + (t == 0 ? P - 1 : t - 1)) * n];
                    const int8_t *nxt = &spins[(r * P + (t + 1) % P) * n];
                    for (std::size_t i = 0; i < n; ++i) {
                        const int8_t s = cur[i];
                        const int nn = static_cast<int>(prv[i]) + static_cast<int>(nxt[i]);
                        const double dcl = -2.0 * static_cast<double>(s) * f[i];
                        const double dtr = 2.0 * jp * static_cast<double>(s) * static_cast<double>(nn);
                        const double de = std::fma(bs, dcl, dtr);  // :115
                        if (de <= 0.0 || uni(g) < std::exp(-de)) {
                            cur[i] = static_cast<int8_t>(-s);
                            m.update(f, cur, i, s);
                        }
                    }
                }
            }
            const double inv = 1.0 / static_cast<double>(n * P);
            for (std::size_t r = 0; r < R; ++r) {
                long long sum = 0;
                for (std::size_t k = 0; k < P * n; ++k) sum += spins[r * P * n + k];
                mvals[r] = static_cast<double>(sum) * inv;
            }
        };
        auto fresh = [&]() {  // random worldline, cached fields, reseeded cell rngs
            random_spins(spins, rng);
            for (std::size_t k = 0; k < R * P; ++k) m.fields(&spins[k * n], &fields[k * n]);
            for (auto &g : crng) g.seed(rng());
        };

        const std::size_t SP = static_cast<std::size_t>(scan_points);
        for (std::size_t q = 0; q < SP; ++q) {  // pilot scan :232-281
            const double s_q = s_lo + (s_hi - s_lo) * static_cast<double>(q) / static_cast<double>(SP - 1);
            const double g_q = s2g(s_q);
            const double jp = trotter_coupling(beta, g_q, P);
            fresh();
            double m2 = 0.0, ma = 0.0;
            std::size_t cnt = 0;
            const double bs = beta / static_cast<double>(P);
            for (int64_t k = 0; k < scan_burn + scan_sweeps; ++k) {
                sweep(bs, jp);
                if (k >= scan_burn)
                    // :253. The reference build vectorizes this loop four replicas at a time:
                    // those products are rounded before the (still sequential) sum, and only the
                    // R % 4 tail is contracted to fma.
                    for (std::size_t j = 0; j < R; ++j) {
                        const double mv = mvals[j];
                        m2 = (j < (R / 4) * 4) ? m2 + mv * mv : std::fma(mv, mv, m2);
                        ma += std::fabs(mv);
                        ++cnt;
                    }
            }
            double chi = 0.0;
            if (cnt >= 2) {
                const double m2m = m2 / static_cast<double>(cnt), mam = ma / static_cast<double>(cnt);
                chi = std::max(static_cast<double>(n * P) * std::fma(-mam, mam, m2m), 0.0);  // :265
            }
            scan_s[q] = s_q;
            scan_gamma[q] = g_q;
            scan_chi[q] = chi;
        }
        const std::size_t peak = static_cast<std::size_t>(std::max_element(scan_chi, scan_chi + SP) - scan_chi);
        scalars[0] = scan_s[peak];
        scalars[1] = scan_gamma[peak];
        scalars[2] = trotter_coupling(beta, scan_gamma[peak], P);
        const double mx = scan_chi[peak];
        const bool degenerate = !(mx > 0.0) || !std::isfinite(mx);
        const double floor = degenerate ? 1.0 : std::max(chi_floor_fraction * mx, 1e-12);
        scalars[3] = floor;
        scalars[4] = A0;

        auto weight = [&](double s) {  // :318-330
            double chi;
            if (s <= scan_s[0]) chi = scan_chi[0];
            else if (s >= scan_s[SP - 1]) chi = scan_chi[SP - 1];
            else {
                std::size_t hi = 1;
                while (hi < SP && scan_s[hi] < s) ++hi;
                const std::size_t lo = hi - 1;
                const double t = (s - scan_s[lo]) / (scan_s[hi] - scan_s[lo]);
                chi = std::fma(t, scan_chi[hi] - scan_chi[lo], scan_chi[lo]);  // :327
            }
            return std::max(chi, floor);
        };
        const std::size_t NS = static_cast<std::size_t>(num_steps);
        std::size_t n1 = 0;
        if (beta_ramp_fraction > 0.0) {
            n1 = static_cast<std::size_t>(beta_ramp_fraction * static_cast<double>(NS));
            if (n1 > NS - 2) n1 = NS - 2;
        }
        const std::size_t n2 = NS - n1;
        const std::size_t GRID = 4096;
        const double hs = (s_hi - s_lo) / static_cast<double>(GRID);
        std::vector<double> W(GRID + 1, 0.0);
        double pw = weight(s_lo);
        for (std::size_t k = 1; k <= GRID; ++k) {  // trapezoid cumulative weight :340-349
            const double sk = std::fma(static_cast<double>(k), hs, s_lo);
            const double cw = weight(sk);
            W[k] = std::fma(0.5 * (pw + cw), hs, W[k - 1]);  // :346
            pw = cw;
        }
        const double Wt = W[GRID];
        for (std::size_t k = 0; k < NS; ++k) {
            beta_sched[k] = beta;
            gamma_sched[k] = gamma_start;
            s_sched[k] = s_lo;
        }
        for (std::size_t k = 0; k < n1; ++k) {  // phase-1 beta ramp :355-360
            const double fr = n1 > 1 ? static_cast<double>(k) / static_cast<double>(n1 - 1) : 1.0;
            beta_sched[k] = std::fma(beta - beta_ramp_start, fr, beta_ramp_start);  // :359
        }
        std::size_t ki = 0;
        for (std::size_t t = 0; t < n2; ++t) {  // inverse-CDF targets :362-376
            const double target = Wt > 0.0 ? (static_cast<double>(t) / static_cast<double>(n2 - 1)) * Wt : 0.0;
            while (ki < GRID && W[ki + 1] < target) ++ki;
            const double w0 = W[ki], w1 = W[ki + 1];
            const double fr = w1 > w0 ? (target - w0) / (w1 - w0) : 0.0;
            const double s_t = std::fma(static_cast<double>(ki) + fr, hs, s_lo);
            s_sched[n1 + t] = s_t;
            gamma_sched[n1 + t] = std::min(std::max(s2g(s_t), gamma_end), gamma_start);
        }

        fresh();  // production on a fresh worldline :394-405
        double best_e = std::numeric_limits<double>::infinity();
        std::vector<int8_t> best(n, 1);
        for (std::size_t k = 0; k < NS; ++k) {
            const double bs = beta_sched[k] / static_cast<double>(P);
            const double jp = trotter_coupling(beta_sched[k], gamma_sched[k], P);
            for (int64_t q = 0; q < sweeps_per_step; ++q) sweep(bs, jp);
            double avg = 0.0;
            for (std::size_t c = 0; c < R * P; ++c) {
                const double e = m.energy(&spins[c * n]);
                avg += e;
                if (e < best_e) { best_e = e; std::memcpy(best.data(), &spins[c * n], n); }
            }
            trace[k] = avg / static_cast<double>(R * P);
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
    });
}

// annealer.cpp:33-75 with an observer that is also a SweepObserver (the trajectory is unchanged)
int orc_sa_observe_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
                       uint64_t seed, int8_t *best_state, double *best_energy, double *trace, double *step_mag,
                       double *sweep_energies, int8_t *sweep_states) {
    return guard([&] {
        if (steps <= 0) throw std::invalid_argument("Schedule must contain at least one beta.");
        if (sweeps_per_beta <= 0) throw std::invalid_argument("sweeps_per_beta must be > 0.");
        const Model m = build(p);
        const std::size_t n = m.n;
        Rng rng(seed);
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        std::vector<int8_t> s(n);
        random_spins(s, rng);
        double e = m.energy(s.data());
        std::vector<int8_t> best = s;
        double best_e = e;
        std::size_t k = 0;
        for (int64_t step = 0; step < steps; ++step) {
            const double beta = betas[step];
            for (int64_t sw = 0; sw < sweeps_per_beta; ++sw) {
                for (std::size_t i = 0; i < n; ++i) {
                    const double d = m.delta(s.data(), i);
                    if (d <= 0.0 || uni(rng) < std::exp(-beta * d)) {
                        s[i] = static_cast<int8_t>(-s[i]);
                        e += d;
                        if (e < best_e) {
                            best_e = e;
                            best = s;
                        }
                    }
                }
                sweep_energies[k] = e;  // :63-65
                std::memcpy(sweep_states + k * n, s.data(), n);
                ++k;
            }
            trace[step] = e;
            double sum = 0.0;  // metrics.hpp magnetization: double sum, one division
            for (std::size_t i = 0; i < n; ++i) sum += static_cast<double>(s[i]);
            step_mag[step] = sum / static_cast<double>(n);
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
    });
}

// sqa_annealer.cpp:87-477 with an observer; tracing selects the sequential path (:317-453)
int orc_sqa_observe_run(const orc_problem *p, int64_t steps, const double *betas, const double *gammas,
                        int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                        int64_t worldline_sweeps, int64_t cluster_sweeps, int32_t tracing, uint64_t seed,
                        int8_t *best_state, double *best_energy, double *trace, double *step_mag,
                        int64_t *sweep_meta, double *sweep_avg, double *sweep_rep, int8_t *sweep_states) {
    return guard([&] {
        if (steps <= 0) throw std::invalid_argument("SQA schedule must contain betas.");
        if (sweeps_per_beta == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
            throw std::invalid_argument(
                "At least one of sweeps_per_beta, worldline_sweeps, or cluster_sweeps must be > 0.");
        const Model m = build(p);
        const std::size_t n = m.n, R = static_cast<std::size_t>(replicas), P = static_cast<std::size_t>(trotter_slices);
        Rng rng(seed);
        std::vector<int8_t> spins(R * P * n);
        random_spins(spins, rng);
        std::vector<Rng> rrng(R);
        for (auto &g : rrng) g.seed(rng());
        std::vector<double> fields(R * P * n);
        for (std::size_t k = 0; k < R * P; ++k) m.fields(&spins[k * n], &fields[k * n]);
        std::vector<std::vector<std::size_t>> order(R, std::vector<std::size_t>(n));
        for (auto &o : order) std::iota(o.begin(), o.end(), std::size_t{0});
        std::vector<std::vector<char>> incl(R, std::vector<char>(P));
        std::vector<std::vector<std::size_t>> stk(R);
        std::vector<std::size_t> ident(n);
        std::iota(ident.begin(), ident.end(), std::size_t{0});
        std::vector<double> rep(R);
        auto avg_energy = [&]() {  // compute_avg_energy (:112-125)
            double total = 0.0;
            for (std::size_t r = 0; r < R; ++r) {
                double sum = 0.0;
                for (std::size_t t = 0; t < P; ++t) sum += m.energy(&spins[(r * P + t) * n]);
                rep[r] = sum / static_cast<double>(P);
                total += sum;
            }
            return total / static_cast<double>(R * P);
        };
        double best_e = std::numeric_limits<double>::infinity();
        std::vector<int8_t> best(n, 1);
        std::size_t k = 0;
        for (int64_t step = 0; step < steps; ++step) {
            const double beta = betas[step], gamma = gammas[step];
            const double jp = trotter_coupling(beta, gamma, P);
            const double bs = beta / static_cast<double>(P);
            const double join = 1.0 - std::exp(-2.0 * jp);
            if (!tracing) {
#pragma omp parallel for if (R > 1) schedule(static)
                for (long long rr = 0; rr < static_cast<long long>(R); ++rr) {
                    const std::size_t r = static_cast<std::size_t>(rr);
                    SqaReplica rep_{&m, n, P, &spins[r * P * n], &fields[r * P * n], &order[r], &incl[r], &stk[r]};
                    for (int64_t q = 0; q < sweeps_per_beta; ++q) rep_.slice_sweep(rrng[r], bs, jp);
                    for (int64_t q = 0; q < cluster_sweeps; ++q) rep_.cluster_sweep(rrng[r], bs, jp, join);
                    for (int64_t q = 0; q < worldline_sweeps; ++q) rep_.worldline_sweep(rrng[r], bs);
                }
            } else {
                std::vector<std::size_t> ord = ident;
                for (std::size_t r = 0; r < R; ++r) {
                    PtReplica tr{&m, n, P, &spins[r * P * n], &ord, &incl[0], &stk[0]};
                    tr.shuffle = false;
                    const struct { int kind, code; int64_t count; } phases[3] = {
                        {0, 0, sweeps_per_beta}, {2, 2, cluster_sweeps}, {1, 1, worldline_sweeps}};
                    for (const auto &ph : phases)
                        for (int64_t q = 0; q < ph.count; ++q) {
                            if (ph.kind == 0) tr.slice_sweep(rng, bs, jp);
                            else if (ph.kind == 2) tr.cluster_sweep(rng, bs, jp, join);
                            else tr.worldline_sweep(rng, bs);
                            const double a = avg_energy();
                            sweep_meta[k * 4] = step;
                            sweep_meta[k * 4 + 1] = static_cast<int64_t>(r);
                            sweep_meta[k * 4 + 2] = q;
                            sweep_meta[k * 4 + 3] = ph.code;
                            sweep_avg[k] = a;
                            std::memcpy(sweep_rep + k * R, rep.data(), R * sizeof(double));
                            std::memcpy(sweep_states + k * R * P * n, spins.data(), R * P * n);
                            ++k;
                        }
                }
            }
            double avg = 0.0;
            for (std::size_t q = 0; q < R * P; ++q) {
                const double e = m.energy(&spins[q * n]);
                avg += e;
                if (e < best_e) {
                    best_e = e;
                    std::memcpy(best.data(), &spins[q * n], n);
                }
            }
            trace[step] = avg / static_cast<double>(R * P);
            double sum = 0.0;
            for (std::size_t q = 0; q < R * P * n; ++q) sum += static_cast<double>(spins[q]);
            step_mag[step] = sum / static_cast<double>(R * P * n);
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
    });
}

// sqa_annealer.cpp:499-952
int orc_sqa_opt_run(const orc_problem *p, int64_t trotter_slices, int64_t replicas, double beta,
                    double j_perp_start, double j_perp_end, double eps_tilde, double alpha,
                    int64_t num_steps, int64_t sweeps_per_step, int64_t worldline_sweeps,
                    int64_t cluster_sweeps, int64_t calib_probes, int64_t calib_sweeps,
                    double beta_ramp_fraction, double beta_ramp_start, uint64_t seed,
                    int8_t *best_state, double *best_energy, double *energy_trace,
                    double *j_perp_trace, double *chi_trace, double *scalars) {
    return guard([&] {
        if (num_steps <= 0) throw std::invalid_argument("num_steps must be > 0.");
        if (sweeps_per_step == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
            throw std::invalid_argument(
                "At least one of sweeps_per_step, worldline_sweeps, or cluster_sweeps must be > 0.");
        if (j_perp_end <= j_perp_start) throw std::invalid_argument("j_perp_end must be > j_perp_start.");
        const bool calibrate = eps_tilde <= 0.0;
        const Model m = build(p);
        const std::size_t n = m.n, P = static_cast<std::size_t>(trotter_slices), R = static_cast<std::size_t>(replicas);
        const std::size_t steps = static_cast<std::size_t>(num_steps);
        const double nM = static_cast<double>(n * P);
        double floor_chi = 1e-4 * nM;
        Rng rng(seed);
        std::vector<int8_t> spins(R * P * n);
        random_spins(spins, rng);  // :534
        std::vector<Rng> rrng(R);
        for (auto &g : rrng) g.seed(rng());  // :541-544
        std::vector<double> fields(R * P * n);
        for (std::size_t k = 0; k < R * P; ++k) m.fields(&spins[k * n], &fields[k * n]);
        std::vector<std::vector<std::size_t>> order(R, std::vector<std::size_t>(n));
        for (auto &o : order) std::iota(o.begin(), o.end(), std::size_t{0});
        std::vector<std::vector<char>> incl(R, std::vector<char>(P));
        std::vector<std::vector<std::size_t>> stk(R);
        std::vector<std::vector<double>> samples(R);

        // one step of every replica's sweeps, one B sample per slice and cluster sweep (:578-726)
        auto sweep = [&](double jp, double bs, int64_t nm, int64_t nc, int64_t nw) {
            const double join = 1.0 - std::exp(-2.0 * jp);
#pragma omp parallel for if (R > 1) schedule(static)
            for (long long rr = 0; rr < static_cast<long long>(R); ++rr) {
                const std::size_t r = static_cast<std::size_t>(rr);
                SqaReplica rep{&m, n, P, &spins[r * P * n], &fields[r * P * n], &order[r], &incl[r], &stk[r]};
                samples[r].clear();
                for (int64_t k = 0; k < nm; ++k) {
                    rep.slice_sweep(rrng[r], bs, jp);
                    samples[r].push_back(bond_sum(rep.spins, P, n));
                }
                for (int64_t k = 0; k < nc; ++k) {
                    rep.cluster_sweep(rrng[r], bs, jp, join);
                    samples[r].push_back(bond_sum(rep.spins, P, n));
                }
                for (int64_t k = 0; k < nw; ++k) rep.worldline_sweep(rrng[r], bs);
            }
        };
        auto chi_raw = [&]() {  // :729-757 with the single-replica guard
            double var = var_b(samples, nM, bond_sum(spins.data(), P, n));
            if (R == 1 && !samples[0].empty()) {
                const double b_avg = samples[0].back() / nM;
                var = std::max(var, nM * std::fma(-b_avg, b_avg, 1.0));
            }
            return var;
        };
        double best_e = std::numeric_limits<double>::infinity();
        std::vector<int8_t> best(n, 1);
        std::size_t rec = 0;
        auto record = [&](double jp, double chi) {  // :761-776
            double avg = 0.0;
            for (std::size_t k = 0; k < R * P; ++k) {
                const double e = m.energy(&spins[k * n]);
                avg += e;
                if (e < best_e) {
                    best_e = e;
                    std::memcpy(best.data(), &spins[k * n], n);
                }
            }
            energy_trace[rec] = avg / static_cast<double>(R * P);
            j_perp_trace[rec] = jp;
            chi_trace[rec] = chi;
            ++rec;
        };
        const double bs_cold = beta / static_cast<double>(P);
        std::size_t n1 = 0;
        if (beta_ramp_fraction > 0.0) {  // :784-789
            n1 = static_cast<std::size_t>(beta_ramp_fraction * static_cast<double>(steps));
            if (n1 >= steps) n1 = steps - 1;
        }
        const std::size_t n2 = steps - n1;
        for (std::size_t k = 0; k < n1; ++k) {  // :801-813
            const double frac = n1 > 1 ? static_cast<double>(k) / static_cast<double>(n1 - 1) : 1.0;
            const double beta_t = std::fma(beta - beta_ramp_start, frac, beta_ramp_start);
            sweep(j_perp_start, beta_t / static_cast<double>(P), sweeps_per_step, cluster_sweeps, worldline_sweeps);
            record(j_perp_start, std::max(chi_raw(), 0.0));
        }
        std::vector<double> pJ, pchi;
        double eps = eps_tilde;
        if (calibrate) {  // :820-876
            const std::size_t M = std::max<std::size_t>(static_cast<std::size_t>(calib_probes), 2);
            const int64_t cs = std::max<int64_t>(calib_sweeps, 1);
            const int64_t pcl = cluster_sweeps > 0 ? std::max<int64_t>(cluster_sweeps, 1) : 0;
            const std::vector<int8_t> saved_s = spins;
            const std::vector<double> saved_f = fields;
            const double dJ = j_perp_end - j_perp_start;
            const double dJp = dJ / static_cast<double>(M);
            pJ.resize(M);
            pchi.assign(M, 0.0);
            double integral = 0.0;
            for (std::size_t q = 0; q < M; ++q) {
                const double Jm = std::fma(static_cast<double>(q) + 0.5, dJp, j_perp_start);
                sweep(Jm, bs_cold, cs, pcl, 0);
                double chi = chi_raw();
                if (!(chi > 0.0)) chi = 0.0;
                pJ[q] = Jm;
                pchi[q] = chi;
                integral = std::fma(std::pow(std::max(chi, 1e-12), alpha), dJp, integral);
            }
            spins = saved_s;
            fields = saved_f;
            eps = integral / static_cast<double>(n2);
            if (!(eps > 0.0) || !std::isfinite(eps)) {
                eps = std::max(dJ, 0.0) / static_cast<double>(steps);
                if (!(eps > 0.0)) eps = 1e-9;
            }
            floor_chi = std::max(1e-4 * *std::max_element(pchi.begin(), pchi.end()), 1e-12);
        }
        std::vector<double> traj;
        if (calibrate && !pJ.empty() && j_perp_end > j_perp_start)
            traj = trajectory(pJ, pchi, j_perp_start, j_perp_end, floor_chi, alpha, n2);
        double j = j_perp_start;
        for (std::size_t k = 0; k < n2; ++k) {  // :925-949
            if (!traj.empty()) j = traj[k];
            sweep(j, bs_cold, sweeps_per_step, cluster_sweeps, worldline_sweeps);
            const double chi_online = chi_raw();
            const double chi_prof = profile_at(pJ, pchi, j);
            const double raw = chi_prof >= 0.0 ? chi_prof : chi_online;
            const double drive = std::max(raw, floor_chi);
            const double dj = eps * std::pow(drive, -alpha);
            record(j, drive);
            if (traj.empty()) j = std::min(j + dj, j_perp_end);
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
        scalars[0] = eps;
        scalars[1] = j_perp_start;
        scalars[2] = j_perp_end;
        scalars[3] = j;
    });
}

// sqa_parallel_tempering.cpp:348-882
int orc_sqapt_opt_run(const orc_problem *p, int64_t rungs, const double *betas, const double *gammas,
                      int64_t trotter_slices, int64_t num_steps, int64_t sweeps_per_step,
                      int64_t worldline_sweeps, double eps_tilde, double alpha, double j_perp_end,
                      int64_t cluster_sweeps, int64_t swap_interval, int64_t continuous_time_slices,
                      int64_t calib_probes, int64_t calib_sweeps, uint64_t seed, int8_t *final_states,
                      double *final_energies, int8_t *best_state, double *best_energy,
                      double *avg_trace, double *swap_trace, double *j_perp_trace, double *chi_trace,
                      double *scalars) {
    return guard([&] {
        if (rungs < 2) throw std::invalid_argument("SQA parallel tempering requires >=2 beta/gamma pairs.");
        if (num_steps <= 0) throw std::invalid_argument("num_steps must be > 0.");
        if (swap_interval <= 0) throw std::invalid_argument("swap_interval must be > 0.");
        if (sweeps_per_step == 0 && worldline_sweeps == 0 && cluster_sweeps == 0)
            throw std::invalid_argument("At least one of sweeps_per_step/worldline_sweeps/cluster_sweeps must be > 0.");
        const bool calibrate = eps_tilde <= 0.0;
        const Model m = build(p);
        const std::size_t n = m.n, R = static_cast<std::size_t>(rungs), steps = static_cast<std::size_t>(num_steps);
        const std::size_t P = (continuous_time_slices > 0 && continuous_time_slices > trotter_slices)
                                  ? static_cast<std::size_t>(continuous_time_slices)
                                  : static_cast<std::size_t>(trotter_slices);
        const double nM = static_cast<double>(n * P);
        const double j0 = trotter_coupling(betas[0], gammas[0], P);  // :375-377
        double j = j0, j1 = j_perp_end;
        if (j1 <= 0.0) j1 = trotter_coupling(betas[R - 1], gammas[R - 1], P);
        if (j1 <= j) j1 = j + 1e6;
        Rng rng(seed);
        std::vector<std::vector<int8_t>> st(R, std::vector<int8_t>(P * n));
        for (auto &s : st) random_spins(s, rng);  // :406-409
        std::vector<Rng> rrng(R);
        for (auto &g : rrng) g.seed(rng());  // :411-412
        std::vector<std::vector<std::size_t>> order(R, std::vector<std::size_t>(n));
        for (auto &o : order) std::iota(o.begin(), o.end(), std::size_t{0});
        std::vector<std::vector<char>> incl(R, std::vector<char>(P));
        std::vector<std::vector<std::size_t>> stk(R);
        std::vector<std::vector<double>> samples(R);
        std::vector<double> bs(R);
        for (std::size_t r = 0; r < R; ++r) bs[r] = betas[r] / static_cast<double>(P);

        auto sweep = [&](std::vector<std::vector<int8_t>> &S, std::vector<Rng> &G, double jp, int64_t nm, int64_t nc,
                         int64_t nw) {  // :454-592
            const double join = 1.0 - std::exp(-2.0 * jp);
#pragma omp parallel for if (R > 1) schedule(static)
            for (long long rl = 0; rl < static_cast<long long>(R); ++rl) {
                const std::size_t r = static_cast<std::size_t>(rl);
                PtReplica rep{&m, n, P, S[r].data(), &order[r], &incl[r], &stk[r]};
                samples[r].clear();
                for (int64_t k = 0; k < nm; ++k) {
                    rep.slice_sweep(G[r], bs[r], jp);
                    samples[r].push_back(bond_sum(rep.spins, P, n));
                }
                for (int64_t k = 0; k < nc; ++k) {
                    rep.cluster_sweep(G[r], bs[r], jp, join);
                    samples[r].push_back(bond_sum(rep.spins, P, n));
                }
                for (int64_t k = 0; k < nw; ++k) rep.worldline_sweep(G[r], bs[r]);
            }
        };
        auto proj = [&](const std::vector<int8_t> &s, std::size_t *bt) {  // :442-451
            double b = std::numeric_limits<double>::infinity();
            std::size_t t0 = 0;
            for (std::size_t t = 0; t < P; ++t) {
                const double e = m.energy(&s[t * n]);
                if (e < b) { b = e; t0 = t; }
            }
            *bt = t0;
            return b;
        };
        auto csum = [&](const std::vector<int8_t> &s) {  // :434-439
            double c = 0.0;
            for (std::size_t t = 0; t < P; ++t) c += m.energy(&s[t * n]);
            return c;
        };
        double best_e = std::numeric_limits<double>::infinity();
        std::vector<int8_t> best(n, 1);
        std::vector<double> en(R);
        auto take_best = [&]() {
            for (std::size_t r = 0; r < R; ++r) {
                std::size_t bt;
                en[r] = proj(st[r], &bt);
                if (en[r] < best_e) { best_e = en[r]; std::memcpy(best.data(), &st[r][bt * n], n); }
            }
        };
        take_best();  // :624-631
        double floor_chi = 1e-4 * nM, eps = eps_tilde;
        std::vector<double> pJ, pchi;
        if (calibrate) {  // :652-711
            const std::size_t M = std::max<std::size_t>(static_cast<std::size_t>(calib_probes), 2);
            const int64_t cs = std::max<int64_t>(calib_sweeps, 1);
            const int64_t pcl = cluster_sweeps > 0 ? std::max<int64_t>(cluster_sweeps, 1) : 0;
            // scratch states and engines; the visit orders are the run's own, shuffled in place
            std::vector<std::vector<int8_t>> scratch = st;
            std::vector<Rng> srng(R);
            for (auto &g : srng) g.seed(rng());
            const double dJ = j1 - j0;
            const double dJp = dJ / static_cast<double>(M);
            pJ.resize(M);
            pchi.assign(M, 0.0);
            double integral = 0.0;
            for (std::size_t q = 0; q < M; ++q) {
                const double Jm = std::fma(static_cast<double>(q) + 0.5, dJp, j0);
                sweep(scratch, srng, Jm, cs, pcl, 0);
                double chi = var_b(samples, nM, bond_sum(scratch[0].data(), P, n));
                if (!(chi > 0.0)) chi = 0.0;
                pJ[q] = Jm;
                pchi[q] = chi;
                integral = std::fma(std::pow(std::max(chi, 1e-12), alpha), dJp, integral);
            }
            eps = integral / static_cast<double>(steps);
            if (!(eps > 0.0) || !std::isfinite(eps)) {
                eps = std::max(dJ, 0.0) / static_cast<double>(steps);
                if (!(eps > 0.0)) eps = 1e-9;
            }
            floor_chi = std::max(1e-4 * *std::max_element(pchi.begin(), pchi.end()), 1e-12);
        }
        std::vector<double> traj;
        if (calibrate && !pJ.empty() && j1 > j0) traj = trajectory(pJ, pchi, j0, j1, floor_chi, alpha, steps);
        for (std::size_t step = 0; step < steps; ++step) {  // :784-872
            if (!traj.empty()) j = traj[step];
            sweep(st, rrng, j, sweeps_per_step, cluster_sweeps, worldline_sweeps);
            take_best();
            const double chi_online = var_b(samples, nM, bond_sum(st[0].data(), P, n));
            const double chi_prof = profile_at(pJ, pchi, j);
            const double raw = chi_prof >= 0.0 ? chi_prof : chi_online;
            const double drive = std::max(raw, floor_chi);
            const double dj = eps * std::pow(drive, -alpha);
            j_perp_trace[step] = j;
            chi_trace[step] = drive;
            if (traj.empty()) j = std::min(j + dj, j1);
            double acc = 0.0, att = 0.0;
            if ((step + 1) % static_cast<std::size_t>(swap_interval) == 0) {
                std::uniform_real_distribution<double> uni(0.0, 1.0);
                for (std::size_t r = 0; r + 1 < R; ++r) {
                    const double c0 = csum(st[r]), c1 = csum(st[r + 1]);
                    const double delta = (bs[r] - bs[r + 1]) * (c1 - c0);
                    att += 1.0;
                    if (delta <= 0.0 || uni(rng) < std::exp(-delta)) {
                        std::swap(st[r], st[r + 1]);
                        std::swap(en[r], en[r + 1]);
                        acc += 1.0;
                    }
                }
            }
            double avg = 0.0;
            for (double e : en) avg += e;
            avg_trace[step] = avg / static_cast<double>(R);
            swap_trace[step] = att > 0.0 ? acc / att : 0.0;
        }
        for (std::size_t r = 0; r < R; ++r) {
            std::size_t bt;
            final_energies[r] = proj(st[r], &bt);
            std::memcpy(final_states + r * n, &st[r][bt * n], n);
        }
        std::memcpy(best_state, best.data(), n);
        *best_energy = best_e;
        scalars[0] = eps;
        scalars[1] = j0;
        scalars[2] = j1;
        scalars[3] = j;
    });
}

}  // extern "C"
