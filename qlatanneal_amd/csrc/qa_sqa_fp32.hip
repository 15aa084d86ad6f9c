// Production SQA sweep (FP32 local fields, Philox-4x32-10), sqa_annealer.cpp:87-477 semantics.
//
// Layout. A workgroup owns a group of rg whole replicas: G = rg*P chains, chain c = (replica
// rr = c / P, slice t = c % P). The group's spins S[G][n] (int8) and FP32 fields F[G][n] stay in
// LDS for the whole run when they fit (one persistent launch, no host round trips); otherwise
// they live in global memory and the same code runs against L2.
//
// Lockstep. All chains of a group share one random visit order per sweep (the reference shares
// one order among a replica's slices, sqa_annealer.cpp:181), and the slices of a replica walk it
// as a skewed wavefront (slice t one position behind slice t-1), which reproduces the reference's
// slice-after-slice sweep order while every chain advances each step. Spins updated together
// never interact, so every step is a product of exact single-spin Metropolis updates.
//
// Per step: lanes = chains decide (dE = (beta/P)(-2 s f) + 2 J_perp s (s_prev + s_next),
// accept if dE <= 0 or u < exp(-dE)); then the workgroup applies f += J[i,:] * (-2 s_old) for the
// flipped chains only. Step tails compute slice energies E = c + 1/2 sum_i s_i (h_i + f_i) in FP64
// from the FP32 fields (direct sum for HUBO), keep each group's first strict minimum, and write
// per-group partial sums of the energy trace; the host reduces groups in replica order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "qa_internal.hpp"
#include "qa_rng.hpp"

namespace qa {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxChains = 64;

enum : uint32_t { TAG_INIT = 1, TAG_ORDER = 2, TAG_SLICE = 3, TAG_WORLD = 4, TAG_CLUSTER = 5 };

struct SqaArgs {
    ProblemView pv;
    int P, R, rg, steps, spb, wl, cl;
    const float *bs;       // [steps] beta / P
    const float *jp;       // [steps] J_perp
    const float *join;     // [steps] 1 - exp(-2 J_perp)
    uint32_t k0, k1;       // Philox key
    int8_t *gS;            // [R*P*n] final spins (and storage when not in LDS)
    float *gF;             // [R*P*n] fields storage when not in LDS
    double *trace_part;    // [steps][ngroups]
    double *best_e;        // [ngroups]
    int *best_step;        // [ngroups]
    int *best_chain;       // [ngroups]
    int8_t *best_state;    // [ngroups][n]
    unsigned long long *accepted;  // [ngroups]
    int ldf;               // row stride of F (LDS: n+1)
};

template <bool kLds>
__global__ void __launch_bounds__(kThreads) sqa_fp32_kernel(SqaArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int n = a.pv.n, P = a.P;
    const int r0 = g * a.rg;
    const int nrep = min(a.rg, a.R - r0);
    const int G = nrep * P;
    const int ldf = a.ldf;

    // LDS carve-up
    unsigned char *p = smem;
    float *F;
    int8_t *S;
    if constexpr (kLds) {
        F = reinterpret_cast<float *>(p);
        p += static_cast<size_t>(a.rg) * P * ldf * sizeof(float);
        S = reinterpret_cast<int8_t *>(p);
        p += ((static_cast<size_t>(a.rg) * P * n + 15) / 16) * 16;
    } else {
        F = a.gF + static_cast<int64_t>(r0) * P * n;
        S = a.gS + static_cast<int64_t>(r0) * P * n;
    }
    int *order = reinterpret_cast<int *>(p);
    p += static_cast<size_t>(n) * sizeof(int);
    float *DS = reinterpret_cast<float *>(p);
    p += kMaxChains * sizeof(float);
    int *ROW = reinterpret_cast<int *>(p);
    p += kMaxChains * sizeof(int);
    double *ered = reinterpret_cast<double *>(p);
    p += kMaxChains * sizeof(double);
    float *wred = reinterpret_cast<float *>(p);
    p += kMaxChains * sizeof(float);
    unsigned long long *cmask = reinterpret_cast<unsigned long long *>(p);
    p += kMaxChains * sizeof(unsigned long long);
    int *rflag = reinterpret_cast<int *>(p);
    p += kMaxChains * sizeof(int);
    unsigned long long *flipmask = reinterpret_cast<unsigned long long *>(p);
    p += 16;
    int *scal = reinterpret_cast<int *>(p);  // [0]: best chain this step

    const uint32_t cbase = static_cast<uint32_t>(r0 * P);  // global chain id of chain 0

    // ---- initial spins (SQAState::random analogue) and fields ------------------------------
    for (int e = tid; e < G * n; e += kThreads) {
        const int c = e / n, i = e - c * n;
        const Philox4 r = philox4x32_10(cbase + c, static_cast<uint32_t>(i), 0u, TAG_INIT, a.k0, a.k1);
        S[static_cast<int64_t>(c) * n + i] = (r.x & 1u) ? 1 : -1;
    }
    __syncthreads();
    for (int e = tid; e < G * n; e += kThreads) {
        const int c = e / n, i = e - c * n;
        const int8_t *sc = S + static_cast<int64_t>(c) * n;
        double acc = a.pv.h32[i];
        if (a.pv.kind == QA_KIND_DENSE) {
            for (int j = 0; j < n; ++j) acc += static_cast<double>(a.pv.J32[static_cast<int64_t>(j) * n + i]) * sc[j];
        } else if (a.pv.kind == QA_KIND_SPARSE) {
            for (int q = a.pv.adj_ptr[i]; q < a.pv.adj_ptr[i + 1]; ++q)
                acc += static_cast<double>(a.pv.adj_val32[q]) * sc[a.pv.adj_idx[q]];
        } else {
            for (int q = a.pv.inc_ptr[i]; q < a.pv.inc_ptr[i + 1]; ++q) {
                const int t = a.pv.inc_idx[q];
                double prod = a.pv.t_coef32[t];
                for (int k = a.pv.t_off[t]; k < a.pv.t_off[t + 1]; ++k)
                    if (a.pv.t_vars[k] != i) prod *= sc[a.pv.t_vars[k]];
                acc += prod;
            }
        }
        F[static_cast<int64_t>(c) * ldf + i] = static_cast<float>(acc);
    }
    double gbest = (tid == 0) ? a.best_e[g] : 0.0;
    unsigned long long nacc = 0;
    __syncthreads();

    const bool decider = tid < G;
    const int c_me = tid;
    const int rr_me = c_me / P, t_me = c_me - (c_me / P) * P;
    const int cprev = rr_me * P + (t_me == 0 ? P - 1 : t_me - 1);
    const int cnext = rr_me * P + (t_me + 1 == P ? 0 : t_me + 1);

    // Apply the field updates of the chains flagged in flipmask (DS/ROW hold ds and spin index).
    auto apply_updates = [&]() {
        unsigned long long m = *flipmask;
        if (m == 0ull) return;
        if (a.pv.kind == QA_KIND_DENSE) {
            while (m) {
                const int c = __ffsll(static_cast<long long>(m)) - 1;
                m &= m - 1ull;
                const float ds = DS[c];
                const float *row = a.pv.J32 + static_cast<int64_t>(ROW[c]) * n;
                float *fc = F + static_cast<int64_t>(c) * ldf;
                for (int j = tid; j < n; j += kThreads) fc[j] = fmaf(ds, row[j], fc[j]);
            }
        } else if (tid < G && ((m >> tid) & 1ull)) {
            const int c = tid, i = ROW[c];
            float *fc = F + static_cast<int64_t>(c) * ldf;
            const int8_t *sc = S + static_cast<int64_t>(c) * n;
            if (a.pv.kind == QA_KIND_SPARSE) {
                const float ds = DS[c];
                for (int q = a.pv.adj_ptr[i]; q < a.pv.adj_ptr[i + 1]; ++q)
                    fc[a.pv.adj_idx[q]] += a.pv.adj_val32[q] * ds;
            } else {
                for (int q = a.pv.inc_ptr[i]; q < a.pv.inc_ptr[i + 1]; ++q) {
                    const int t = a.pv.inc_idx[q];
                    float prod = a.pv.t_coef32[t];
                    for (int k = a.pv.t_off[t]; k < a.pv.t_off[t + 1]; ++k) prod *= sc[a.pv.t_vars[k]];
                    for (int k = a.pv.t_off[t]; k < a.pv.t_off[t + 1]; ++k) {
                        const int v = a.pv.t_vars[k];
                        if (v != i) fc[v] += 2.0f * prod * sc[v];
                    }
                }
            }
        }
    };

    // Random permutation of 0..n-1 for this sweep (Fisher-Yates on one lane, Philox draws).
    auto make_order = [&](uint32_t sweep_id) {
        if (tid == 0) {
            for (int i = 0; i < n; ++i) order[i] = i;
            for (int i = n - 1; i > 0; i -= 4) {
                const Philox4 r = philox4x32_10(cbase, sweep_id, static_cast<uint32_t>(i), TAG_ORDER, a.k0, a.k1);
                const uint32_t d[4] = {r.x, r.y, r.z, r.w};
                for (int q = 0; q < 4 && i - q > 0; ++q) {
                    const int ii = i - q;
                    const int j = static_cast<int>(below(d[q], static_cast<uint32_t>(ii + 1)));
                    const int tmp = order[ii];
                    order[ii] = order[j];
                    order[j] = tmp;
                }
            }
        }
        __syncthreads();
    };

    uint32_t sweep_id = 0;
    for (int step = 0; step < a.steps; ++step) {
        const float bs = a.bs[step], jp = a.jp[step], join = a.join[step];
        // ---------------- slice sweeps (sqa_annealer.cpp:180-209) --------------------------
        // Skewed wavefront: at step tau, slice t works on order position k = tau - t. Slice t-1
        // has already passed position k (so s_{t-1}[i_k] is its value after this sweep) and slice
        // t+1 has not reached it (s_{t+1}[i_k] is still the value before this sweep) -- exactly
        // what the reference's slice-by-slice loop sees, including the ring closure where slice
        // P-1 reads slice 0's swept value. Simultaneous updates always touch distinct spin
        // indices, so they never interact. A sweep takes n + P - 1 steps.
        for (int sw = 0; sw < a.spb; ++sw, ++sweep_id) {
            make_order(sweep_id);
            Philox4 rnd{0, 0, 0, 0};
            const int nsteps = n + P - 1;
            for (int tau = 0; tau < nsteps; ++tau) {
                if (tid < kMaxChains) {
                    const int k = tau - t_me;
                    const bool active = decider && k >= 0 && k < n;
                    bool acc = false;
                    if ((tau & 3) == 0)
                        rnd = philox4x32_10(cbase + c_me, sweep_id, static_cast<uint32_t>(tau >> 2), TAG_SLICE,
                                            a.k0, a.k1);
                    if (active) {
                        const int i = order[k];
                        int8_t *sc = S + static_cast<int64_t>(c_me) * n;
                        const int8_t s = sc[i];
                        const float f = F[static_cast<int64_t>(c_me) * ldf + i];
                        const int nb = S[static_cast<int64_t>(cprev) * n + i] + S[static_cast<int64_t>(cnext) * n + i];
                        const float fs = static_cast<float>(s);
                        const float de = fmaf(bs, -2.0f * fs * f, 2.0f * jp * fs * static_cast<float>(nb));
                        const uint32_t q = tau & 3;
                        const uint32_t word = q == 0 ? rnd.x : q == 1 ? rnd.y : q == 2 ? rnd.z : rnd.w;
                        acc = de <= 0.0f || u01(word) < __expf(-de);
                        DS[c_me] = -2.0f * fs;
                        ROW[c_me] = i;
                        if (acc) sc[i] = static_cast<int8_t>(-s);
                    }
                    const unsigned long long m = __ballot(acc);
                    if (tid == 0) *flipmask = m;
                    nacc += acc ? 1ull : 0ull;
                }
                __syncthreads();
                apply_updates();
                __syncthreads();
            }
        }
        // ---------------- cluster sweeps (sqa_annealer.cpp:215-288) ------------------------
        for (int sw = 0; sw < a.cl; ++sw, ++sweep_id) {
            make_order(sweep_id);
            for (int k = 0; k < n; ++k) {
                const int i = order[k];
                if (decider && t_me == 0) {
                    // leader lane of replica rr_me grows the imaginary-time cluster (LIFO, prev
                    // before next), membership as a bit mask over the replica's P <= 64 slices
                    const int base = rr_me * P;
                    uint32_t draw = 0;
                    auto next_u = [&]() {
                        const Philox4 r = philox4x32_10(cbase + base, sweep_id, static_cast<uint32_t>(k),
                                                        TAG_CLUSTER | (draw << 8), a.k0, a.k1);
                        ++draw;
                        return r.x;
                    };
                    const int seed = static_cast<int>(below(next_u(), static_cast<uint32_t>(P)));
                    const int8_t s0 = S[static_cast<int64_t>(base + seed) * n + i];
                    unsigned long long inc = 1ull << seed, stack = 1ull << seed;
                    while (stack) {
                        const int t = 63 - __clzll(static_cast<long long>(stack));
                        stack &= ~(1ull << t);
                        const int pv = t == 0 ? P - 1 : t - 1, nx = t + 1 == P ? 0 : t + 1;
                        if (!((inc >> pv) & 1ull) && S[static_cast<int64_t>(base + pv) * n + i] == s0 &&
                            u01(next_u()) < join) { inc |= 1ull << pv; stack |= 1ull << pv; }
                        if (!((inc >> nx) & 1ull) && S[static_cast<int64_t>(base + nx) * n + i] == s0 &&
                            u01(next_u()) < join) { inc |= 1ull << nx; stack |= 1ull << nx; }
                    }
                    float dcl = 0.0f, dtm = 0.0f;
                    for (int t = 0; t < P; ++t) {
                        if (!((inc >> t) & 1ull)) continue;
                        dcl += -2.0f * s0 * F[static_cast<int64_t>(base + t) * ldf + i];
                        const int pv = t == 0 ? P - 1 : t - 1, nx = t + 1 == P ? 0 : t + 1;
                        if (!((inc >> pv) & 1ull)) dtm += 2.0f * jp * s0 * S[static_cast<int64_t>(base + pv) * n + i];
                        if (!((inc >> nx) & 1ull)) dtm += 2.0f * jp * s0 * S[static_cast<int64_t>(base + nx) * n + i];
                    }
                    const float de = fmaf(bs, dcl, dtm);
                    const bool ok = de <= 0.0f || u01(next_u()) < __expf(-de);
                    cmask[rr_me] = ok ? inc : 0ull;
                }
                __syncthreads();
                bool acc = false;
                if (decider) {
                    acc = (cmask[rr_me] >> t_me) & 1ull;
                    int8_t *sc = S + static_cast<int64_t>(c_me) * n;
                    const int8_t s = sc[i];
                    DS[c_me] = -2.0f * s;
                    ROW[c_me] = i;
                    if (acc) sc[i] = static_cast<int8_t>(-s);
                }
                if (tid < kMaxChains) {
                    const unsigned long long m = __ballot(acc);
                    if (tid == 0) *flipmask = m;
                }
                __syncthreads();
                apply_updates();
                __syncthreads();
            }
        }
        // ---------------- worldline sweeps (sqa_annealer.cpp:294-315) -----------------------
        for (int sw = 0; sw < a.wl; ++sw, ++sweep_id) {
            make_order(sweep_id);
            for (int k = 0; k < n; ++k) {
                const int i = order[k];
                if (decider) wred[c_me] = -2.0f * S[static_cast<int64_t>(c_me) * n + i] * F[static_cast<int64_t>(c_me) * ldf + i];
                __syncthreads();
                if (decider && t_me == 0) {
                    float dcl = 0.0f;
                    for (int t = 0; t < P; ++t) dcl += wred[rr_me * P + t];
                    const float de = bs * dcl;
                    const Philox4 r = philox4x32_10(cbase + rr_me * P, sweep_id, static_cast<uint32_t>(k),
                                                    TAG_WORLD, a.k0, a.k1);
                    rflag[rr_me] = (de <= 0.0f || u01(r.x) < __expf(-de)) ? 1 : 0;
                }
                __syncthreads();
                bool acc = false;
                if (decider) {
                    acc = rflag[rr_me] != 0;
                    int8_t *sc = S + static_cast<int64_t>(c_me) * n;
                    const int8_t s = sc[i];
                    DS[c_me] = -2.0f * s;
                    ROW[c_me] = i;
                    if (acc) sc[i] = static_cast<int8_t>(-s);
                }
                if (tid < kMaxChains) {
                    const unsigned long long m = __ballot(acc);
                    if (tid == 0) *flipmask = m;
                }
                __syncthreads();
                apply_updates();
                __syncthreads();
            }
        }
        // ---------------- step tail: slice energies, group minimum, trace partial ----------
        {
            const int wave = tid >> 6, lane = tid & 63;
            for (int c = wave; c < G; c += kThreads / 64) {
                const int8_t *sc = S + static_cast<int64_t>(c) * n;
                const float *fc = F + static_cast<int64_t>(c) * ldf;
                double acc = 0.0;
                if (a.pv.kind != QA_KIND_HUBO) {
                    for (int i = lane; i < n; i += 64)
                        acc += static_cast<double>(sc[i]) * (static_cast<double>(a.pv.h32[i]) + static_cast<double>(fc[i]));
                    acc *= 0.5;
                } else {
                    for (int i = lane; i < n; i += 64) acc += static_cast<double>(a.pv.h32[i]) * sc[i];
                    for (int t = lane; t < a.pv.nterms; t += 64) {
                        double prod = a.pv.t_coef32[t];
                        for (int q = a.pv.t_off[t]; q < a.pv.t_off[t + 1]; ++q) prod *= sc[a.pv.t_vars[q]];
                        acc += prod;
                    }
                }
                for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
                if (lane == 0) ered[c] = a.pv.c + acc;
            }
            __syncthreads();
            if (tid == 0) {
                double sum = 0.0, mn = gbest;
                int arg = -1;
                for (int c = 0; c < G; ++c) {
                    sum += ered[c];
                    if (ered[c] < mn) { mn = ered[c]; arg = c; }
                }
                a.trace_part[static_cast<int64_t>(step) * gridDim.x + g] = sum;
                scal[0] = arg;
                if (arg >= 0) {
                    gbest = mn;
                    a.best_e[g] = mn;
                    a.best_step[g] = step;
                    a.best_chain[g] = r0 * P + arg;
                }
            }
            __syncthreads();
            const int arg = scal[0];
            if (arg >= 0)
                for (int i = tid; i < n; i += kThreads)
                    a.best_state[static_cast<int64_t>(g) * n + i] = S[static_cast<int64_t>(arg) * n + i];
            __syncthreads();
        }
    }
    // final spins out (for the caller's state readback / diagnostics)
    if constexpr (kLds) {
        for (int e = tid; e < G * n; e += kThreads) a.gS[static_cast<int64_t>(r0) * P * n + e] = S[e];
    }
    if (tid < kMaxChains) {
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        if (tid == 0) a.accepted[g] = nacc;
    }
}

}  // namespace

void sqa_fp32(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res) {
    const int64_t n = p->n, R = prm->replicas, P = prm->trotter_slices;
    if (P > kMaxChains)
        throw Error{QA_EUNSUPPORTED, "FP32 mode runs at most 64 Trotter slices per replica in this build"};
    if (p->kind == QA_KIND_DENSE && !p->symmetric)
        throw Error{QA_EUNSUPPORTED, "FP32 mode needs a symmetric J (the reference's field updates assume it); "
                                     "use the FP64 validation mode for an asymmetric J"};
    QA_REQUIRE(n < (1 << 24), "problem too large");
    // replicas per group: fill a 64-lane wave when there are enough replicas for every CU
    int64_t rg = std::max<int64_t>(1, kMaxChains / P);
    while (rg > 1 && (R + rg - 1) / rg < ctx->num_cus) --rg;
    const int ldf = static_cast<int>(n + 1);
    auto lds_bytes = [&](int64_t r, bool lds) {
        size_t b = static_cast<size_t>(n) * 4 + kMaxChains * (4 + 4 + 8 + 4 + 8 + 4) + 16 + 16;
        if (lds) b += static_cast<size_t>(r * P) * ldf * 4 + ((static_cast<size_t>(r * P * n) + 15) / 16) * 16;
        return b;
    };
    const size_t lds_cap = 160 * 1024;
    while (rg > 1 && lds_bytes(rg, true) > lds_cap) --rg;
    const bool use_lds = lds_bytes(rg, true) <= lds_cap;
    const int64_t ngroups = (R + rg - 1) / rg;
    const size_t shmem = lds_bytes(rg, use_lds);
    QA_REQUIRE(shmem <= lds_cap, "problem too large for the FP32 sweep kernel's LDS layout");

    hipStream_t s = ctx->stream;
    std::vector<float> hbs(prm->steps), hjp(prm->steps), hjoin(prm->steps);
    for (int64_t k = 0; k < prm->steps; ++k) {
        const double jp = trotter_coupling(prm->betas[k], prm->gammas[k], P);
        hbs[k] = static_cast<float>(prm->betas[k] / static_cast<double>(P));
        hjp[k] = static_cast<float>(jp);
        hjoin[k] = static_cast<float>(1.0 - std::exp(-2.0 * jp));
    }
    DevBuf<float> dbs(prm->steps), djp(prm->steps), djoin(prm->steps);
    dbs.upload(hbs.data(), prm->steps, s);
    djp.upload(hjp.data(), prm->steps, s);
    djoin.upload(hjoin.data(), prm->steps, s);
    DevBuf<int8_t> gS(R * P * n), bstate(ngroups * n);
    DevBuf<float> gF(use_lds ? 1 : R * P * n);
    DevBuf<double> tpart(prm->steps * ngroups), be(ngroups);
    DevBuf<int> bstep(ngroups), bchain(ngroups);
    DevBuf<unsigned long long> accd(ngroups);
    std::vector<double> inf(ngroups, std::numeric_limits<double>::infinity());
    be.upload(inf.data(), ngroups, s);
    QA_HIP(hipMemsetAsync(bchain.ptr, 0xff, ngroups * sizeof(int), s));

    SqaArgs a{};
    a.pv = view_of(p);
    a.P = static_cast<int>(P);
    a.R = static_cast<int>(R);
    a.rg = static_cast<int>(rg);
    a.steps = static_cast<int>(prm->steps);
    a.spb = static_cast<int>(prm->sweeps_per_beta);
    a.wl = static_cast<int>(prm->worldline_sweeps);
    a.cl = static_cast<int>(prm->cluster_sweeps);
    a.bs = dbs.ptr;
    a.jp = djp.ptr;
    a.join = djoin.ptr;
    a.k0 = static_cast<uint32_t>(prm->seed);
    a.k1 = static_cast<uint32_t>(prm->seed >> 32);
    a.gS = gS.ptr;
    a.gF = gF.ptr;
    a.trace_part = tpart.ptr;
    a.best_e = be.ptr;
    a.best_step = bstep.ptr;
    a.best_chain = bchain.ptr;
    a.best_state = bstate.ptr;
    a.accepted = accd.ptr;
    a.ldf = use_lds ? ldf : static_cast<int>(n);

    auto kern = use_lds ? sqa_fp32_kernel<true> : sqa_fp32_kernel<false>;
    QA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem)));
    QA_HIP(hipEventRecord(ctx->ev0, s));
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(ngroups)), dim3(kThreads), shmem, s, a);
    QA_HIP(hipGetLastError());
    QA_HIP(hipEventRecord(ctx->ev1, s));

    std::vector<double> hpart(prm->steps * ngroups), hbe(ngroups);
    std::vector<int> hstep(ngroups), hchain(ngroups);
    std::vector<unsigned long long> hacc(ngroups);
    tpart.download(hpart.data(), hpart.size(), s);
    be.download(hbe.data(), ngroups, s);
    bstep.download(hstep.data(), ngroups, s);
    bchain.download(hchain.data(), ngroups, s);
    accd.download(hacc.data(), ngroups, s);
    QA_HIP(hipStreamSynchronize(s));
    float ms = 0;
    QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    for (int64_t k = 0; k < prm->steps; ++k) {
        double sum = 0.0;
        for (int64_t gi = 0; gi < ngroups; ++gi) sum += hpart[k * ngroups + gi];
        res->energy_trace[k] = sum / static_cast<double>(R * P);
    }
    // best over groups: first strict minimum in (step, replica, slice) order, as the reference
    int64_t bg = -1;
    for (int64_t gi = 0; gi < ngroups; ++gi) {
        if (hchain[gi] < 0) continue;
        if (bg < 0 || hbe[gi] < hbe[bg] || (hbe[gi] == hbe[bg] && hstep[gi] < hstep[bg])) bg = gi;
    }
    QA_REQUIRE(bg >= 0, "no finite slice energy was produced");
    QA_HIP(hipMemcpyAsync(res->best_state, bstate.ptr + bg * n, n, hipMemcpyDeviceToHost, s));
    DevBuf<double> ex(1);
    device_energies(p, bstate.ptr + bg * n, 1, ex.ptr, s);
    ex.download(res->best_energy, 1, s);
    QA_HIP(hipStreamSynchronize(s));
    if (res->kernel_ms) *res->kernel_ms = ms;
    if (res->accept_rate) {
        unsigned long long tot = 0;
        for (auto v : hacc) tot += v;
        const double att = static_cast<double>(prm->steps) * prm->sweeps_per_beta * R * P * n;
        *res->accept_rate = att > 0 ? static_cast<double>(tot) / att : 0.0;
    }
}

}  // namespace qa
