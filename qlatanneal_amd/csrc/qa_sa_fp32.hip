// Production SA (FP32 local fields, Philox-4x32-10), annealer.cpp:33-75 semantics for `reads`
// independent chains in one launch.
//
// A workgroup owns up to 64 reads (lane = read for the Metropolis decision). Spins and FP32
// fields of its reads stay in LDS when they fit, else in global memory. Sites are visited in
// index order 0..n-1 as in the reference; after each step the whole workgroup applies
// f += J[i,:] * (-2 s_old) for the reads that flipped (all of them use row i, loaded once).
// Energies: E starts from an FP64 evaluation (fields accumulated in FP64) and accumulates the
// accepted dE; the best state is copied after every improving accepted flip, as in the
// reference, and the returned best energy is recomputed in FP64 from the returned state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "qa_internal.hpp"
#include "qa_rng.hpp"

namespace qa {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxReads = 64;
enum : uint32_t { TAG_SA_INIT = 11, TAG_SA_STEP = 12 };

struct SaArgs {
    ProblemView pv;
    int reads, steps, sweeps;
    const float *betas;     // [steps]
    uint32_t k0, k1;
    int8_t *gS;             // [reads][n] (storage when not in LDS)
    float *gF;              // [reads][n]
    int8_t *best;           // [reads][n]
    double *best_e;         // [reads]  (running-energy value at the best state)
    double *trace;          // [reads][steps]
    int ldf;
    // chained runs (optional)
    const float *read_betas;     // [reads]: per-read beta at every step (null: betas[step])
    const int8_t *init_states;   // [reads][n]: start states (null: Philox random)
    const double *init_e;        // [reads]: running energy at the start (null: FP64 energy)
    uint32_t ctr_base;           // Philox counter word 2 (distinct streams for chained runs)
    int8_t *final_states;        // [reads][n] (null: not written)
    double *mag;                 // [reads][steps]: mean spin at each step end
    double *btrace;              // [reads][steps+1]: best running energy at the start and each step end
};

template <bool kLds>
__global__ void __launch_bounds__(kThreads) sa_fp32_kernel(SaArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const int n = a.pv.n;
    const int r0 = blockIdx.x * kMaxReads;
    const int G = min(kMaxReads, a.reads - r0);
    const int ldf = a.ldf;
    unsigned char *p = smem;
    float *F;
    int8_t *S;
    if constexpr (kLds) {
        F = reinterpret_cast<float *>(p);
        p += static_cast<size_t>(kMaxReads) * ldf * sizeof(float);
        S = reinterpret_cast<int8_t *>(p);
        p += ((static_cast<size_t>(kMaxReads) * n + 15) / 16) * 16;
    } else {
        F = a.gF + static_cast<int64_t>(r0) * n;
        S = a.gS + static_cast<int64_t>(r0) * n;
    }
    float *DS = reinterpret_cast<float *>(p);
    p += kMaxReads * sizeof(float);
    double *E = reinterpret_cast<double *>(p);
    p += kMaxReads * sizeof(double);
    unsigned long long *masks = reinterpret_cast<unsigned long long *>(p);  // [0] flips, [1] improved

    for (int e = tid; e < G * n; e += kThreads) {
        const int c = e / n, i = e - c * n;
        if (a.init_states) {
            S[static_cast<int64_t>(c) * n + i] = a.init_states[static_cast<int64_t>(r0) * n + e];
        } else {
            const Philox4 r = philox4x32_10(static_cast<uint32_t>(r0 + c), static_cast<uint32_t>(i), 0u, TAG_SA_INIT,
                                            a.k0, a.k1);
            S[static_cast<int64_t>(c) * n + i] = (r.x & 1u) ? 1 : -1;
        }
    }
    if (tid < kMaxReads) E[tid] = 0.0;
    __syncthreads();
    // fields in FP64 (stored as FP32) and the exact starting energy
    for (int c = 0; c < G; ++c) {
        const int8_t *sc = S + static_cast<int64_t>(c) * n;
        double part = 0.0;
        for (int i = tid; i < n; i += kThreads) {
            double acc = a.pv.h64[i];
            if (a.pv.kind == QA_KIND_DENSE) {
                const double *row = a.pv.J64 + static_cast<int64_t>(i) * n;
                for (int j = 0; j < n; ++j)
                    if (j != i) acc += row[j] * sc[j];
            } else if (a.pv.kind == QA_KIND_SPARSE) {
                for (int q = a.pv.adj_ptr[i]; q < a.pv.adj_ptr[i + 1]; ++q) acc += a.pv.adj_val64[q] * sc[a.pv.adj_idx[q]];
            } else {
                for (int q = a.pv.inc_ptr[i]; q < a.pv.inc_ptr[i + 1]; ++q) {
                    const int t = a.pv.inc_idx[q];
                    double prod = a.pv.t_coef64[t];
                    for (int k = a.pv.t_off[t]; k < a.pv.t_off[t + 1]; ++k)
                        if (a.pv.t_vars[k] != i) prod *= sc[a.pv.t_vars[k]];
                    acc += prod;
                }
            }
            F[static_cast<int64_t>(c) * ldf + i] = static_cast<float>(acc);
            if (a.pv.kind != QA_KIND_HUBO) part += 0.5 * sc[i] * (a.pv.h64[i] + acc);
        }
        if (a.pv.kind == QA_KIND_HUBO) {
            for (int i = tid; i < n; i += kThreads) part += a.pv.h64[i] * sc[i];
            for (int t = tid; t < a.pv.nterms; t += kThreads) {
                double prod = a.pv.t_coef64[t];
                for (int q = a.pv.t_off[t]; q < a.pv.t_off[t + 1]; ++q) prod *= sc[a.pv.t_vars[q]];
                part += prod;
            }
        }
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
        if ((tid & 63) == 0) atomicAdd(&E[c], part);
    }
    __syncthreads();
    double ecur = 0.0, ebest = 0.0;
    if (tid < G) {
        ecur = a.init_e ? a.init_e[r0 + tid] : a.pv.c + E[tid];
        ebest = ecur;
        if (a.btrace) a.btrace[static_cast<int64_t>(r0 + tid) * (a.steps + 1)] = ebest;
    }
    for (int i = tid; i < G * n; i += kThreads) a.best[static_cast<int64_t>(r0) * n + i] = S[(i / n) * static_cast<int64_t>(n) + (i % n)];
    __syncthreads();

    uint32_t ctr = 0;
    for (int step = 0; step < a.steps; ++step) {
        const float beta = (a.read_betas && tid < G) ? a.read_betas[r0 + tid] : a.betas[step];
        for (int sw = 0; sw < a.sweeps; ++sw) {
            Philox4 rnd{0, 0, 0, 0};
            for (int i = 0; i < n; ++i, ++ctr) {
                if (tid < kMaxReads) {
                    bool acc = false, imp = false;
                    // a sweep may start inside a four-word block (n % 4 != 0): draw the block then too
                    if ((ctr & 3) == 0 || i == 0)
                        rnd = philox4x32_10(static_cast<uint32_t>(r0 + tid), ctr >> 2, a.ctr_base, TAG_SA_STEP, a.k0, a.k1);
                    if (tid < G) {
                        int8_t *sc = S + static_cast<int64_t>(tid) * n;
                        const float fs = static_cast<float>(sc[i]);
                        const float de = -2.0f * fs * F[static_cast<int64_t>(tid) * ldf + i];
                        const uint32_t q = ctr & 3;
                        const uint32_t w = q == 0 ? rnd.x : q == 1 ? rnd.y : q == 2 ? rnd.z : rnd.w;
                        acc = de <= 0.0f || u01(w) < __expf(-beta * de);
                        DS[tid] = -2.0f * fs;
                        if (acc) {
                            sc[i] = static_cast<int8_t>(-sc[i]);
                            ecur += static_cast<double>(de);
                            if (ecur < ebest) {
                                ebest = ecur;
                                imp = true;
                            }
                        }
                    }
                    const unsigned long long fm = __ballot(acc), im = __ballot(imp);
                    if (tid == 0) {
                        masks[0] = fm;
                        masks[1] = im;
                    }
                }
                __syncthreads();
                unsigned long long m = masks[0];
                if (m) {
                    if (a.pv.kind == QA_KIND_DENSE) {
                        const float *row = a.pv.J32 + static_cast<int64_t>(i) * n;
                        for (int j = tid; j < n; j += kThreads) {
                            const float rj = row[j];
                            unsigned long long mm = m;
                            while (mm) {
                                const int c = __ffsll(static_cast<long long>(mm)) - 1;
                                mm &= mm - 1ull;
                                float *fc = F + static_cast<int64_t>(c) * ldf;
                                fc[j] = fmaf(DS[c], rj, fc[j]);
                            }
                        }
                    } else if (tid < G && ((m >> tid) & 1ull)) {
                        float *fc = F + static_cast<int64_t>(tid) * ldf;
                        const int8_t *sc = S + static_cast<int64_t>(tid) * n;
                        if (a.pv.kind == QA_KIND_SPARSE) {
                            for (int q = a.pv.adj_ptr[i]; q < a.pv.adj_ptr[i + 1]; ++q)
                                fc[a.pv.adj_idx[q]] += a.pv.adj_val32[q] * DS[tid];
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
                    unsigned long long im = masks[1];
                    while (im) {
                        const int c = __ffsll(static_cast<long long>(im)) - 1;
                        im &= im - 1ull;
                        for (int j = tid; j < n; j += kThreads)
                            a.best[static_cast<int64_t>(r0 + c) * n + j] = S[static_cast<int64_t>(c) * n + j];
                    }
                }
                __syncthreads();
            }
        }
        if (tid < G) {
            a.trace[static_cast<int64_t>(r0 + tid) * a.steps + step] = ecur;
            if (a.btrace) a.btrace[static_cast<int64_t>(r0 + tid) * (a.steps + 1) + step + 1] = ebest;
            if (a.mag) {  // magnetization (metrics.hpp:11-20): integer sum, one division
                const int8_t *sc = S + static_cast<int64_t>(tid) * n;
                int sum = 0;
                for (int i = 0; i < n; ++i) sum += sc[i];
                a.mag[static_cast<int64_t>(r0 + tid) * a.steps + step] = static_cast<double>(sum) / static_cast<double>(n);
            }
        }
    }
    if (tid < G) {
        a.best_e[r0 + tid] = ebest;
    }
    if (a.final_states)
        for (int e = tid; e < G * n; e += kThreads) a.final_states[static_cast<int64_t>(r0) * n + e] = S[e];
}

// ---- small dense problems: one wavefront per read -------------------------------------------
// For n <= 128 the block kernel above spends a site step on two workgroup barriers for one
// attempt per read. Here a wavefront owns one read: site i's spin and FP32 field live in lane
// i % 64 (register slot i / 64), the Metropolis decision is taken uniformly by every lane from the
// value read through v_readlane, and an accepted flip updates every lane's field with one fmaf
// from the J row staged in LDS (the row of the next site is loaded before its decision). Random
// words, dE, the acceptance test and the field update are those of sa_fp32_kernel (same Philox
// counters), so both kernels produce the same spin trajectory for a given key.
constexpr int kWaveMaxN = 128;

__device__ inline float readlane_f32(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// NQ: register slots per lane (site i lives in lane i % 64, slot i / 64); n <= 64 NQ. The site
// loop is straight-line: every update is a select on the uniform acceptance, so no branch is
// taken per site.
template <int NQ>
__global__ void __launch_bounds__(256) sa_wave_kernel(SaArgs a) {
    extern __shared__ __align__(16) float Jl[];  // [n][n] FP32 couplings, zero diagonal
    const int n = a.pv.n, lane = threadIdx.x & 63;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) Jl[e] = a.pv.J32[e];
    __syncthreads();
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= a.reads) return;  // wave-uniform; no barrier follows
    const int64_t rn = static_cast<int64_t>(r) * n;
    float sp[NQ], fl[NQ], bsp[NQ], u[NQ];
    bool valid[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int i = lane + 64 * q;
        valid[q] = i < n;
        sp[q] = 1.0f;
        if (valid[q]) {
            if (a.init_states) {
                sp[q] = static_cast<float>(a.init_states[rn + i]);
            } else {
                const Philox4 w = philox4x32_10(static_cast<uint32_t>(r), static_cast<uint32_t>(i), 0u, TAG_SA_INIT,
                                                a.k0, a.k1);
                sp[q] = (w.x & 1u) ? 1.0f : -1.0f;
            }
        }
    }
    // FP64 fields in the block kernel's order (j ascending, j != i), rounded once; the starting
    // energy c + sum_i s_i (h_i + f_i) / 2 as a wave reduction
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int i = lane + 64 * q;
        double acc = valid[q] ? a.pv.h64[i] : 0.0;
        for (int j = 0; j < n; ++j) {
            float sj = 0.0f;
#pragma unroll
            for (int qq = 0; qq < NQ; ++qq)
                if ((j >> 6) == qq) sj = readlane_f32(sp[qq], j & 63);
            if (valid[q] && j != i) acc += a.pv.J64[static_cast<int64_t>(i) * n + j] * static_cast<double>(sj);
        }
        fl[q] = static_cast<float>(acc);
        if (valid[q]) part += 0.5 * static_cast<double>(sp[q]) * (a.pv.h64[i] + acc);
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    double ecur = a.init_e ? a.init_e[r] : a.pv.c + part;
    double ebest = ecur;
#pragma unroll
    for (int q = 0; q < NQ; ++q) bsp[q] = sp[q];
    if (a.btrace && lane == 0) a.btrace[static_cast<int64_t>(r) * (a.steps + 1)] = ebest;
    uint32_t ctr = 0;
    for (int step = 0; step < a.steps; ++step) {
        const float beta = a.read_betas ? a.read_betas[r] : a.betas[step];
        for (int sw = 0; sw < a.sweeps; ++sw, ctr += static_cast<uint32_t>(n)) {
            // this sweep's uniforms, one per site, from the same Philox words as the block kernel
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const uint32_t c = ctr + static_cast<uint32_t>(lane + 64 * q);
                const Philox4 w = philox4x32_10(static_cast<uint32_t>(r), c >> 2, a.ctr_base, TAG_SA_STEP, a.k0, a.k1);
                const uint32_t k = c & 3u;
                u[q] = u01(k == 0 ? w.x : k == 1 ? w.y : k == 2 ? w.z : w.w);
            }
            float jr[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) jr[q] = Jl[min(lane + 64 * q, n - 1)];  // row 0
            for (int i = 0; i < n; ++i) {
                const int li = i & 63;
                float fi = 0.0f, si = 0.0f, ui = 0.0f;
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    if (NQ == 1 || (i >> 6) == q) {
                        fi = readlane_f32(fl[q], li);
                        si = readlane_f32(sp[q], li);
                        ui = readlane_f32(u[q], li);
                    }
                float rj[NQ];
                const int nxt = min(i + 1, n - 1);  // next site's row, loaded before this decision resolves
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    rj[q] = jr[q];
                    jr[q] = Jl[nxt * n + min(lane + 64 * q, n - 1)];
                }
                const float de = -2.0f * si * fi;
                const bool acc = de <= 0.0f || ui < __expf(-beta * de);
                const float ds = -2.0f * si;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const float upd = fmaf(ds, rj[q], fl[q]);
                    fl[q] = (acc && valid[q]) ? upd : fl[q];
                    sp[q] = (acc && (NQ == 1 || (i >> 6) == q) && lane == li) ? -sp[q] : sp[q];
                }
                ecur = acc ? ecur + static_cast<double>(de) : ecur;
                const bool imp = acc && ecur < ebest;
                ebest = imp ? ecur : ebest;
#pragma unroll
                for (int q = 0; q < NQ; ++q) bsp[q] = imp ? sp[q] : bsp[q];
            }
        }
        if (lane == 0) a.trace[static_cast<int64_t>(r) * a.steps + step] = ecur;
        if (a.btrace && lane == 0) a.btrace[static_cast<int64_t>(r) * (a.steps + 1) + step + 1] = ebest;
        if (a.mag) {  // magnetization (metrics.hpp:11-20): integer sum, one division
            int sum = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) sum += valid[q] ? static_cast<int>(sp[q]) : 0;
            for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
            if (lane == 0) a.mag[static_cast<int64_t>(r) * a.steps + step] = static_cast<double>(sum) / static_cast<double>(n);
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int i = lane + 64 * q;
        if (valid[q]) {
            a.best[rn + i] = static_cast<int8_t>(bsp[q]);
            if (a.final_states) a.final_states[rn + i] = static_cast<int8_t>(sp[q]);
        }
    }
    if (lane == 0) a.best_e[r] = ebest;
}

}  // namespace

void sa_fp32(qa_ctx *ctx, const qa_problem *p, const qa_sa_params *prm, qa_sa_result *res) {
    if (p->kind == QA_KIND_DENSE && !p->symmetric)
        throw Error{QA_EUNSUPPORTED, "FP32 mode needs a symmetric J; use the FP64 validation mode for an asymmetric J"};
    const int64_t n = p->n, R = prm->reads;
    const int ldf = static_cast<int>(n + 1);
    const size_t small = kMaxReads * (4 + 8) + 32;
    const size_t big = small + static_cast<size_t>(kMaxReads) * ldf * 4 + ((static_cast<size_t>(kMaxReads) * n + 15) / 16) * 16;
    const bool use_lds = big <= 160 * 1024;
    const size_t shmem = use_lds ? big : small;
    hipStream_t s = ctx->stream;
    std::vector<float> hb(prm->steps);
    for (int64_t k = 0; k < prm->steps; ++k) hb[k] = static_cast<float>(prm->betas[k]);
    DevBuf<float> betas(prm->steps), gF(use_lds ? 1 : R * n);
    DevBuf<float> rbeta(prm->read_betas ? R : 0);
    // Final energies are the FP64 energies of the final states, recomputed on the device, so that
    // chained runs (parallel tempering) carry exact energies rather than sums of FP32 deltas.
    const bool want_final = res->final_states || res->final_energies;
    DevBuf<int8_t> istates(prm->init_states ? R * n : 0), fstates(want_final ? R * n : 0);
    DevBuf<double> ienergy(prm->init_energies ? R : 0), fenergy(res->final_energies ? R : 0),
        mag(res->magnetization_trace ? R * prm->steps : 0), btr(res->best_energy_trace ? R * (prm->steps + 1) : 0);
    if (prm->read_betas) {
        std::vector<float> rb(R);
        for (int64_t r = 0; r < R; ++r) rb[r] = static_cast<float>(prm->read_betas[r]);
        rbeta.upload(rb.data(), R, s);
        QA_HIP(hipStreamSynchronize(s));
    }
    istates.upload(prm->init_states, prm->init_states ? R * n : 0, s);
    ienergy.upload(prm->init_energies, prm->init_energies ? R : 0, s);
    DevBuf<int8_t> gS(use_lds ? 1 : R * n), best(R * n);
    DevBuf<double> be(R), tr(R * prm->steps), ex(R);
    betas.upload(hb.data(), prm->steps, s);
    SaArgs a{};
    a.pv = view_of(p);
    a.reads = static_cast<int>(R);
    a.steps = static_cast<int>(prm->steps);
    a.sweeps = static_cast<int>(prm->sweeps_per_beta);
    a.betas = betas.ptr;
    a.k0 = static_cast<uint32_t>(prm->seed);
    a.k1 = static_cast<uint32_t>(prm->seed >> 32);
    a.gS = gS.ptr;
    a.gF = gF.ptr;
    a.best = best.ptr;
    a.best_e = be.ptr;
    a.trace = tr.ptr;
    a.ldf = use_lds ? ldf : static_cast<int>(n);
    a.read_betas = rbeta.ptr;
    a.init_states = istates.ptr;
    a.init_e = ienergy.ptr;
    a.ctr_base = static_cast<uint32_t>(prm->counter_base);
    a.final_states = fstates.ptr;
    a.mag = mag.ptr;
    a.btrace = btr.ptr;
    // small dense problems: one wavefront per read (QANNEAL_SA_KERNEL=block forces the block kernel)
    const char *env = std::getenv("QANNEAL_SA_KERNEL");
    const bool wave = p->kind == QA_KIND_DENSE && n <= kWaveMaxN && !(env && std::string(env) == "block");
    QA_HIP(hipEventRecord(ctx->ev0, s));
    if (wave) {
        const size_t jsh = static_cast<size_t>(n) * n * sizeof(float);
        auto kern = n <= 64 ? sa_wave_kernel<1> : sa_wave_kernel<2>;
        ensure_dynamic_lds(reinterpret_cast<const void *>(kern), jsh);
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>((R + 3) / 4)), dim3(256), jsh, s, a);
    } else {
        auto kern = use_lds ? sa_fp32_kernel<true> : sa_fp32_kernel<false>;
        ensure_dynamic_lds(reinterpret_cast<const void *>(kern), shmem);
        const unsigned grid = static_cast<unsigned>((R + kMaxReads - 1) / kMaxReads);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), shmem, s, a);
    }
    QA_HIP(hipGetLastError());
    QA_HIP(hipEventRecord(ctx->ev1, s));
    device_energies(p, best.ptr, R, ex.ptr, s);
    if (res->final_energies) device_energies(p, fstates.ptr, R, fenergy.ptr, s);
    best.download(res->best_states, R * n, s);
    ex.download(res->best_energies, R, s);
    tr.download(res->energy_trace, R * prm->steps, s);
    if (res->final_states) fstates.download(res->final_states, R * n, s);
    if (res->final_energies) fenergy.download(res->final_energies, R, s);
    if (res->magnetization_trace) mag.download(res->magnetization_trace, R * prm->steps, s);
    if (res->best_energy_trace) btr.download(res->best_energy_trace, R * (prm->steps + 1), s);
    QA_HIP(hipStreamSynchronize(s));
    if (res->uniforms_used)
        for (int64_t r = 0; r < R; ++r) res->uniforms_used[r] = 0;
    if (res->kernel_ms) {
        float ms = 0;
        QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *res->kernel_ms = ms;
    }
}

}  // namespace qa
