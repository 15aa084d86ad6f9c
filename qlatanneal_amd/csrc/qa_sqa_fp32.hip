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
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "qa_internal.hpp"
#include "qa_rng.hpp"

namespace qa {
namespace {

// Pipelined slice sweeps give the waves roles: wave 0 decides (one lane per chain), waves 1-4
// apply field updates, wave 5 stages J rows and shuffles the next visit order. The other sweep
// kinds use all threads alike.
constexpr int kThreads = 384;
constexpr int kUpd = 256;        // update threads (tid 64..319)
constexpr int kHelperWave = 5;
constexpr int kPreH = 32;        // helper registers for one staged row (n <= 64 * kPreH)
constexpr int kMaxChains = 64;  // chains of a group (rg * P) in one wavefront: pipelined sweeps, rg > 1
// Trotter slices per replica of the two-barrier kernel: one thread per chain, one replica per
// group once P > 64 (the reference accepts any count, sqa_annealer.cpp:28-30, 97-99).
constexpr int kMaxSlices = kThreads;

// LDS regions start on 16-byte boundaries (the per-chain double arrays need 8).
__host__ __device__ constexpr size_t lds_align16(size_t b) { return (b + 15) / 16 * 16; }

// Four floats as a plain vector value (16-byte LDS and global accesses).
typedef float f4 __attribute__((ext_vector_type(4)));

// fn(std::integral_constant<int, 0>) ... fn(std::integral_constant<int, N-1>): register arrays
// indexed this way stay in registers.
template <typename Fn, int... Is>
__device__ __forceinline__ void static_for_impl(Fn &&fn, std::integer_sequence<int, Is...>) {
    (fn(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename Fn>
__device__ __forceinline__ void static_for(Fn &&fn) {
    static_for_impl(fn, std::make_integer_sequence<int, N>{});
}

enum : uint32_t { TAG_INIT = 1, TAG_ORDER = 2, TAG_SLICE = 3, TAG_WORLD = 4, TAG_CLUSTER = 5, TAG_CHECKER = 6 };

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
    int ldf;               // row stride of F (LDS layout: n rounded up to 4; global layout: n)
    int ldr;               // row stride of the J-row ring (n rounded up to 4)
    int ring;              // J rows staged in LDS (0: read rows from global memory)
    int pfd;               // prefetch distance of the row ring, in steps
    const int *param_of_rep;  // non-null: replica rep uses bs/jp/join[param_of_rep[rep]] (ladder rungs)
    int resume;            // 1: continue from gS/gF instead of a fresh random state
    int save_fields;       // 1: write F back to gF at the end (LDS layout)
    int tail_mode;         // 0: group minimum + trace partials; 1: per-chain energies and bond sums
    double *chain_e;       // tail_mode 1: [R*P] slice energies at the end of the launch
    int *chain_b;          // tail_mode 1: [R*P] sum_i s_t,i s_t+1,i
    uint32_t sweep_base;   // sweep counter offset of this launch (random-stream continuity)
    int step_base;         // schedule step of this launch's step 0 (recorded in best_step)
    int ngroups;           // replica groups of the run (row stride of trace_part)
    int sweep_kind;        // 0: SQA slice sweeps (wavefront); 1: sqa_chi parity checkerboard
    int mag_burn;          // checkerboard: sweeps before magnetisation is accumulated
    double *mag_m2;        // [R] sum over measured sweeps of m^2 (null: no measurement)
    double *mag_abs;       // [R] sum over measured sweeps of |m|
    // Imaginary-time bond sums B_r = sum_t sum_i s_t,i s_t+1,i (adaptive J_perp schedules): one
    // sample of every replica after each slice sweep and each cluster sweep (worldline flips
    // keep B). Integer-valued, so the double sums below are exact while they stay under 2^53.
    double *bstat;         // non-null: [steps][ngroups][2] sum of B and of B^2 over the group's samples
    double *bend;          // [steps]: B of replica 0 at the end of each step (group 0 writes)
    int hubo_lds;          // HUBO: incidence and term arrays staged in LDS after the scratch
    int hubo_nv;           // HUBO: entries of t_vars (= of inc_idx)
    int adj_lds;           // sparse: adjacency (CSR) staged in LDS after the scratch
    int cmax;              // per-chain scratch entries: 64, or G rounded up to 64 when P > 64
    int s_lds;             // global-state fields with the spins in LDS (dense float4 kernel)
    int la;                // look-ahead slice sweep: its LDS exchange area follows the spins
};


// ---- pieces shared by the sweep kernels -----------------------------------------------------
// Group state: F[G][ldf] fields and S[G][n] spins of chains r0*P .. r0*P + G - 1 (LDS or global).

// LDS carve-up shared by both sweep kernels (sizes as computed by plan_sqa).
struct SqaLds {
    float *F;          // [rg*P][ldf] fields (LDS layout)
    int8_t *S;         // [rg*P][n] spins (LDS layout)
    int *order;        // two visit-order buffers of obstride ints
    int obstride;
    float *Jring;      // [ring][ldf]: J rows of upcoming order positions
    int *RSLOT;        // [cmax] ring slot of each chain's site (two-barrier sweeps)
    float *DS;         // [cmax] -2 s_old of each chain's attempt (two-barrier sweeps)
    int *ROW;          // [cmax] site of each chain's attempt (two-barrier sweeps)
    double *ered;      // [cmax] slice energies of the step tail
    float *wred;       // [cmax] worldline partial sums
    unsigned long long *cmask;  // [cmax] cluster memberships
    int *rflag;        // [cmax] worldline decisions
    unsigned long long *flipmask;  // two-barrier: [w] flips of wave w's chains; pipelined: [b] flips, [2 + b] flips from s = +1
    int *scal;         // [0]: best chain of the step
    float *junk;       // [64][4]: target of stores a lane must not make (never read)
    float *la;         // look-ahead sweep exchange area (kLaBytes)
    const int32_t *inc_ptr, *inc_idx, *t_off, *t_vars;  // HUBO arrays (LDS copies or global)
    const float *t_coef;
    const int32_t *adj_ptr, *adj_idx;                     // sparse CSR (LDS copy or global)
    const float *adj_val;
};

// LDS bytes of a staged sparse adjacency (adj_ptr, adj_idx, adj_val32).
__host__ __device__ constexpr size_t adj_lds_bytes(int n, int nnz) {
    return lds_align16(static_cast<size_t>(n + 1) * 4) + 2 * lds_align16(static_cast<size_t>(nnz) * 4);
}

// LDS bytes of the staged HUBO arrays (inc_ptr, inc_idx, t_off, t_vars, t_coef32).
__host__ __device__ constexpr size_t hubo_lds_bytes(int n, int nterms, int nv) {
    return lds_align16(static_cast<size_t>(n + 1) * 4) + 2 * lds_align16(static_cast<size_t>(nv) * 4) +
           lds_align16(static_cast<size_t>(nterms + 1) * 4) + lds_align16(static_cast<size_t>(nterms) * 4);
}

// Flip-mask words: one per deciding wavefront (cmax / 64), at least the pipelined kernel's four.
// Look-ahead sweep: the helper wave loads a decision's field element kLaLead lockstep steps before
// it (kLaLead - 1 steps for the loads to land), and the decision adds the chain's flips of the
// kLaCorr = kLaLead + 1 steps the element does not hold yet. Exchange area in LDS, in 32-bit words:
// PF [2][64] at 0, JC [kLaCorr][2][64], DS [2][64], ROW [2][64], two 64-bit flip masks, the Philox
// words RW [2][4][64].
constexpr int kLaLead = 3;
constexpr int kLaCorr = kLaLead + 1;
constexpr int kLaJC = 128;
constexpr int kLaDS = kLaJC + kLaCorr * 128;
constexpr int kLaROW = kLaDS + 128;
constexpr int kLaFM = kLaROW + 128;
constexpr int kLaRW = kLaFM + 4;
constexpr size_t kLaBytes = static_cast<size_t>(kLaRW + 512) * 4;

__host__ __device__ constexpr size_t flipmask_bytes(int cmax) {
    return lds_align16(static_cast<size_t>(cmax / 64 > 4 ? cmax / 64 : 4) * 8);
}

// LDS of the per-chain scratch (RSLOT, DS, ROW, ered, wred, cmask, rflag), the flip masks, scal
// and junk, for cmax chain entries.
__host__ __device__ constexpr size_t chain_scratch_bytes(int cmax) {
    return static_cast<size_t>(cmax) * (4 + 4 + 4 + 8 + 4 + 8 + 4) + flipmask_bytes(cmax) + 16 + 64 * 16;
}

__device__ SqaLds carve_lds(const SqaArgs &a, unsigned char *p, bool lds_state) {
    const int n = a.pv.n;
    SqaLds L{};
    if (lds_state) {
        L.F = reinterpret_cast<float *>(p);
        p += lds_align16(static_cast<size_t>(a.rg) * a.P * a.ldf * sizeof(float));
        L.S = reinterpret_cast<int8_t *>(p);
        p += lds_align16(static_cast<size_t>(a.rg) * a.P * n);
    }
    L.order = reinterpret_cast<int *>(p);
    L.obstride = static_cast<int>(lds_align16(static_cast<size_t>(n) * sizeof(int)) / sizeof(int));
    p += 2 * lds_align16(static_cast<size_t>(n) * sizeof(int));
    L.Jring = reinterpret_cast<float *>(p);
    p += lds_align16(static_cast<size_t>(a.ring) * a.ldr * sizeof(float));
    const int cm = a.cmax;
    L.RSLOT = reinterpret_cast<int *>(p);
    p += cm * sizeof(int);
    L.DS = reinterpret_cast<float *>(p);
    p += cm * sizeof(float);
    L.ROW = reinterpret_cast<int *>(p);
    p += cm * sizeof(int);
    L.ered = reinterpret_cast<double *>(p);
    p += cm * sizeof(double);
    L.wred = reinterpret_cast<float *>(p);
    p += cm * sizeof(float);
    L.cmask = reinterpret_cast<unsigned long long *>(p);
    p += cm * sizeof(unsigned long long);
    L.rflag = reinterpret_cast<int *>(p);
    p += cm * sizeof(int);
    L.flipmask = reinterpret_cast<unsigned long long *>(p);
    p += flipmask_bytes(cm);
    L.scal = reinterpret_cast<int *>(p);
    p += 16;
    L.junk = reinterpret_cast<float *>(p);
    p += 64 * 16;
    L.inc_ptr = a.pv.inc_ptr;
    L.inc_idx = a.pv.inc_idx;
    L.t_off = a.pv.t_off;
    L.t_vars = a.pv.t_vars;
    L.t_coef = a.pv.t_coef32;
    L.adj_ptr = a.pv.adj_ptr;
    L.adj_idx = a.pv.adj_idx;
    L.adj_val = a.pv.adj_val32;
    if (a.adj_lds) {  // the plan placed them here, after everything above
        const int nnz = 2 * a.pv.nedges;
        L.adj_ptr = reinterpret_cast<const int32_t *>(p);
        p += lds_align16(static_cast<size_t>(n + 1) * 4);
        L.adj_idx = reinterpret_cast<const int32_t *>(p);
        p += lds_align16(static_cast<size_t>(nnz) * 4);
        L.adj_val = reinterpret_cast<const float *>(p);
    }
    if (!lds_state && a.s_lds) {  // the plan placed them here, after everything above
        L.S = reinterpret_cast<int8_t *>(p);
        p += lds_align16(static_cast<size_t>(a.rg) * a.P * n);
    }
    if (a.la) {  // after the spins
        L.la = reinterpret_cast<float *>(p);
        p += kLaBytes;
    }
    if (a.hubo_lds) {  // the plan placed them here, after everything above
        L.inc_ptr = reinterpret_cast<const int32_t *>(p);
        p += lds_align16(static_cast<size_t>(n + 1) * 4);
        L.inc_idx = reinterpret_cast<const int32_t *>(p);
        p += lds_align16(static_cast<size_t>(a.hubo_nv) * 4);
        L.t_off = reinterpret_cast<const int32_t *>(p);
        p += lds_align16(static_cast<size_t>(a.pv.nterms + 1) * 4);
        L.t_vars = reinterpret_cast<const int32_t *>(p);
        p += lds_align16(static_cast<size_t>(a.hubo_nv) * 4);
        L.t_coef = reinterpret_cast<const float *>(p);
    }
    return L;
}

// Copies the sparse adjacency or the HUBO arrays into their LDS places (called before
// group_init's closing barrier).
__device__ void lists_stage(const SqaArgs &a, const SqaLds &L) {
    const int tid = threadIdx.x, nt = blockDim.x, n = a.pv.n, nterms = a.pv.nterms;
    if (a.adj_lds) {
        for (int e = tid; e <= n; e += nt) const_cast<int32_t *>(L.adj_ptr)[e] = a.pv.adj_ptr[e];
        for (int e = tid; e < 2 * a.pv.nedges; e += nt) {
            const_cast<int32_t *>(L.adj_idx)[e] = a.pv.adj_idx[e];
            const_cast<float *>(L.adj_val)[e] = a.pv.adj_val32[e];
        }
        return;
    }
    for (int e = tid; e <= n; e += nt) const_cast<int32_t *>(L.inc_ptr)[e] = a.pv.inc_ptr[e];
    for (int e = tid; e < a.hubo_nv; e += nt) {
        const_cast<int32_t *>(L.inc_idx)[e] = a.pv.inc_idx[e];
        const_cast<int32_t *>(L.t_vars)[e] = a.pv.t_vars[e];
    }
    for (int e = tid; e <= nterms; e += nt) const_cast<int32_t *>(L.t_off)[e] = a.pv.t_off[e];
    for (int e = tid; e < nterms; e += nt) const_cast<float *>(L.t_coef)[e] = a.pv.t_coef32[e];
}

// Random permutation of 0..n-1 for sweep sweep_id of the group whose first chain is cbase
// (Fisher-Yates by one thread, Philox draws: four per counter block); the caller synchronises.
__device__ void shuffle_order(int *order, int n, uint32_t cbase, uint32_t sweep_id, uint32_t k0, uint32_t k1) {
    for (int i = 0; i < n; ++i) order[i] = i;
    for (int i = n - 1; i > 0; i -= 4) {
        const Philox4 r = philox4x32_10(cbase, sweep_id, static_cast<uint32_t>(i), TAG_ORDER, k0, k1);
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


// Initial spins (SQAState::random analogue) and FP32 fields, or the state of a previous launch.
__device__ void group_init(const SqaArgs &a, float *F, int8_t *S, int ldf, int G, int r0, bool lds_state) {
    const int tid = threadIdx.x, nt = blockDim.x, n = a.pv.n, P = a.P;
    const uint32_t cbase = static_cast<uint32_t>(r0 * P);
    if (a.resume) {
        if (lds_state) {
            for (int e = tid; e < G * n; e += nt) {
                const int c = e / n, i = e - c * n;
                S[e] = a.gS[static_cast<int64_t>(r0) * P * n + e];
                F[static_cast<int64_t>(c) * ldf + i] = a.gF[static_cast<int64_t>(r0) * P * n + e];
            }
        } else if (a.s_lds) {  // fields stay in global memory, spins come into LDS
            for (int e = tid; e < G * n; e += nt) S[e] = a.gS[static_cast<int64_t>(r0) * P * n + e];
        }
    } else {
        for (int e = tid; e < G * n; e += nt) {
            const int c = e / n, i = e - c * n;
            const Philox4 r = philox4x32_10(cbase + c, static_cast<uint32_t>(i), 0u, TAG_INIT, a.k0, a.k1);
            S[static_cast<int64_t>(c) * n + i] = (r.x & 1u) ? 1 : -1;
        }
        __syncthreads();
        for (int e = tid; e < G * n; e += nt) {
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
        // row padding (columns n .. ldf-1) stays zero
        for (int e = tid; e < G * (ldf - n); e += nt) {
            const int c = e / (ldf - n);
            F[static_cast<int64_t>(c) * ldf + n + (e - c * (ldf - n))] = 0.0f;
        }
    }
    __syncthreads();
}

// Step tail: slice energies E = c + 1/2 sum_i s_i (h_i + f_i) in FP64 from the FP32 fields (direct
// sum for HUBO). tail_mode 0 keeps the group's first strict minimum (best_e/best_step/best_chain/
// best_state; gbest lives in thread 0) and writes the group's partial of the energy trace;
// tail_mode 1 writes per-chain energies and bond sums sum_i s_t,i s_t+1,i.
__device__ void group_tail(const SqaArgs &a, const float *F, const int8_t *S, int ldf, int G, int r0, int g,
                           int step, double *ered, int *scal, double &gbest) {
    const int tid = threadIdx.x, nt = blockDim.x, n = a.pv.n, P = a.P;
    const int wave = tid >> 6, lane = tid & 63;
    for (int c = wave; c < G; c += nt / 64) {
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
        if (a.tail_mode == 1) {
            const int t = c % P;
            const int8_t *sn = S + static_cast<int64_t>((c / P) * P + (t + 1 == P ? 0 : t + 1)) * n;
            int b = 0;
            for (int i = lane; i < n; i += 64) b += sc[i] * sn[i];
            for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o);
            if (lane == 0) {
                a.chain_e[static_cast<int64_t>(r0) * P + c] = a.pv.c + acc;
                a.chain_b[static_cast<int64_t>(r0) * P + c] = b;
            }
        }
    }
    __syncthreads();
    if (tid == 0 && a.tail_mode == 0) {
        double sum = 0.0, mn = gbest;
        int arg = -1;
        for (int c = 0; c < G; ++c) {
            sum += ered[c];
            if (ered[c] < mn) { mn = ered[c]; arg = c; }
        }
        a.trace_part[static_cast<int64_t>(step) * a.ngroups + g] = sum;
        scal[0] = arg;
        if (arg >= 0) {
            gbest = mn;
            a.best_e[g] = mn;
            a.best_step[g] = a.step_base + step;
            a.best_chain[g] = r0 * P + arg;
        }
    }
    __syncthreads();
    const int arg = a.tail_mode == 0 ? scal[0] : -1;
    if (arg >= 0)
        for (int i = tid; i < n; i += nt) a.best_state[static_cast<int64_t>(g) * n + i] = S[static_cast<int64_t>(arg) * n + i];
    __syncthreads();
}

// Final spins (and, for a resumed sequence of launches, fields) of an LDS-resident group.
__device__ void group_save(const SqaArgs &a, const float *F, const int8_t *S, int ldf, int G, int r0) {
    const int tid = threadIdx.x, nt = blockDim.x, n = a.pv.n, P = a.P;
    for (int e = tid; e < G * n; e += nt) {
        a.gS[static_cast<int64_t>(r0) * P * n + e] = S[e];
        if (a.save_fields) {
            const int c = e / n, i = e - c * n;
            a.gF[static_cast<int64_t>(r0) * P * n + e] = F[static_cast<int64_t>(c) * ldf + i];
        }
    }
}

// One replica group of the two-barrier sweep (any model, LDS or global state).
// NJ > 0 (global state, dense, n % 4 == 0, n <= 1024 NJ): field updates in float4 groups, NJ per
// thread, with every load of a batch of flipped chains issued before any of its stores, so a batch
// costs one round trip to L2/MALL instead of one per column stride (same fmaf per element).
template <bool kLds, int NJ = 0, bool kSL = false>
__device__ void general_group(const SqaArgs &a, const int g) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const int n = a.pv.n, P = a.P;
    const int r0 = g * a.rg;
    const int nrep = min(a.rg, a.R - r0);
    const int G = nrep * P;
    const int ldf = a.ldf;

    const SqaLds L = carve_lds(a, smem, kLds);
    float *F = kLds ? L.F : a.gF + static_cast<int64_t>(r0) * P * n;
    // kSL: fields in global memory, spins in LDS (copied back at the end)
    int8_t *S = (kLds || kSL) ? L.S : a.gS + static_cast<int64_t>(r0) * P * n;
    int *order = L.order;
    float *Jring = L.Jring;
    int *RSLOT = L.RSLOT;
    float *DS = L.DS;
    int *ROW = L.ROW;
    double *ered = L.ered;
    float *wred = L.wred;
    unsigned long long *cmask = L.cmask;
    int *rflag = L.rflag;
    unsigned long long *flipmask = L.flipmask;
    int *scal = L.scal;

    const uint32_t cbase = static_cast<uint32_t>(r0 * P);  // global chain id of chain 0

    if (a.hubo_lds || a.adj_lds) lists_stage(a, L);  // visible after group_init's closing barrier
    group_init(a, F, S, ldf, G, r0, kLds);
    double gbest = (tid == 0 && a.tail_mode == 0) ? a.best_e[g] : 0.0;
    unsigned long long nacc = 0;

    const bool decider = tid < G;
    const int c_me = tid;
    const int rr_me = c_me / P, t_me = c_me - (c_me / P) * P;
    const int cprev = rr_me * P + (t_me == 0 ? P - 1 : t_me - 1);
    const int cnext = rr_me * P + (t_me + 1 == P ? 0 : t_me + 1);

    // Apply the field updates of the chains flagged in flipmask (DS/ROW hold ds and spin index).
    // Dense: the flipped chains are taken eight at a time with a wave-uniform mask; every LDS
    // load of a batch (ring row value, field) is issued before any of its stores, so one batch
    // costs one LDS round trip instead of one per chain and column.
    // One flip-mask word: chains cw .. cw + 63. skip (global-state slice sweeps): element SKIP[c]
    // of a flipped chain's row keeps its value -- the chain's next decision has loaded it already
    // and applies this flip to it itself (and writes it back).
    int *SKIP = RSLOT;   // the row ring is never used with the state in global memory
    int *SKIP2 = rflag;  // (worldline decisions; free during slice sweeps): kSL write-back element
    // pre(): loads to issue right after the first batch's row loads (global-state slice sweeps:
    // the next decisions' inputs), so both sets are in flight together; pre_done records it.
    auto apply_word = [&](const unsigned long long m_all, const int cw, const bool skip, auto &&pre, bool &pre_done) {
        if (m_all == 0ull) return;
        if constexpr (!kLds && NJ > 0) {
            constexpr int BMAX = NJ >= 8 ? 1 : 8 / NJ;
            uint32_t mlo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all));
            uint32_t mhi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all >> 32));
            unsigned long long m = (static_cast<unsigned long long>(mhi) << 32) | mlo;
            const int n4 = n >> 2;
            f4 *F4g = reinterpret_cast<f4 *>(F);
            // batches of exactly BMAX, ..., 2 or 1 flipped chains: every load of a batch is issued
            // unconditionally (lanes past the row end reread its last group), so the batch's loads
            // are in flight together with no wait between them; only the stores are masked
            auto batch = [&](auto bsz) {
                constexpr int B = decltype(bsz)::value;
                int cs[B], sk[B], sk2[B];
                float dsv[B];
                const f4 *rows[B];
                const f4 *frow[B];
#pragma unroll
                for (int q = 0; q < B; ++q) {
                    const int c = cw + __ffsll(static_cast<long long>(m)) - 1;
                    m &= m - 1ull;
                    cs[q] = c;
                    dsv[q] = DS[c];
                    sk[q] = skip ? SKIP[c] : -1;
                    sk2[q] = (kSL && skip) ? SKIP2[c] : -1;
                    rows[q] = reinterpret_cast<const f4 *>(a.pv.J32 + static_cast<int64_t>(ROW[c]) * n);
                    frow[q] = F4g + static_cast<int64_t>(c) * n4;
                }
                f4 rv[B][NJ], fv[B][NJ];
#pragma unroll
                for (int q = 0; q < B; ++q)
#pragma unroll
                    for (int u = 0; u < NJ; ++u) {
                        const int gi = min(tid + kThreads * u, n4 - 1);
                        rv[q][u] = rows[q][gi];
                        fv[q][u] = frow[q][gi];
                    }
                if (!pre_done) {
                    pre();
                    pre_done = true;
                }
#pragma unroll
                for (int q = 0; q < B; ++q)
#pragma unroll
                    for (int u = 0; u < NJ; ++u) {
                        const int gi = tid + kThreads * u;
                        if (gi < n4) {
                            const f4 x = rv[q][u], y = fv[q][u];
                            const f4 v = f4{fmaf(dsv[q], x.x, y.x), fmaf(dsv[q], x.y, y.y), fmaf(dsv[q], x.z, y.z),
                                            fmaf(dsv[q], x.w, y.w)};
                            if (gi == (sk[q] >> 2) || gi == (sk2[q] >> 2)) {
                                // a group holding a skipped element: the other elements one by one
                                float *d = reinterpret_cast<float *>(F4g + static_cast<int64_t>(cs[q]) * n4 + gi);
                                const int b0 = 4 * gi;
                                if (b0 != sk[q] && b0 != sk2[q]) d[0] = v.x;
                                if (b0 + 1 != sk[q] && b0 + 1 != sk2[q]) d[1] = v.y;
                                if (b0 + 2 != sk[q] && b0 + 2 != sk2[q]) d[2] = v.z;
                                if (b0 + 3 != sk[q] && b0 + 3 != sk2[q]) d[3] = v.w;
                            } else {
                                F4g[static_cast<int64_t>(cs[q]) * n4 + gi] = v;
                            }
                        }
                    }
            };
            while (m) {
                const int cnt = __popcll(m);
                if (BMAX >= 8 && cnt >= 8) batch(std::integral_constant<int, (BMAX >= 8 ? 8 : 1)>{});
                else if (BMAX >= 4 && cnt >= 4) batch(std::integral_constant<int, (BMAX >= 4 ? 4 : 1)>{});
                else if (BMAX >= 2 && cnt >= 2) batch(std::integral_constant<int, (BMAX >= 2 ? 2 : 1)>{});
                else batch(std::integral_constant<int, 1>{});
            }
            return;
        }
        if (a.pv.kind == QA_KIND_DENSE) {
            uint32_t mlo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all));
            uint32_t mhi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all >> 32));
            unsigned long long m = (static_cast<unsigned long long>(mhi) << 32) | mlo;
            // batches of exactly 8, 4, 2 or 1 flipped chains: every load of a column is issued
            // unconditionally, and the J source is one address space per loop, so the loads of a
            // column are in flight together
            auto batch = [&](auto bsz) {
                constexpr int B = decltype(bsz)::value;
                float dsv[B];
                int64_t roff[B], foff[B];
#pragma unroll
                for (int q = 0; q < B; ++q) {
                    const int c = cw + __ffsll(static_cast<long long>(m)) - 1;
                    m &= m - 1ull;
                    dsv[q] = DS[c];
                    roff[q] = static_cast<int64_t>(a.ring ? RSLOT[c] : ROW[c]) * n;
                    foff[q] = static_cast<int64_t>(c) * ldf;
                }
                auto columns = [&](const float *src) {
                    for (int j = tid; j < n; j += kThreads) {
                        float rv[B], fv[B];
#pragma unroll
                        for (int q = 0; q < B; ++q) {
                            rv[q] = src[roff[q] + j];
                            fv[q] = F[foff[q] + j];
                        }
#pragma unroll
                        for (int q = 0; q < B; ++q) F[foff[q] + j] = fmaf(dsv[q], rv[q], fv[q]);
                    }
                };
                if (a.ring) columns(Jring);
                else columns(a.pv.J32);
            };
            while (m) {
                const int cnt = __popcll(m);
                if (cnt >= 8) batch(std::integral_constant<int, 8>{});
                else if (cnt >= 4) batch(std::integral_constant<int, 4>{});
                else if (cnt >= 2) batch(std::integral_constant<int, 2>{});
                else batch(std::integral_constant<int, 1>{});
            }
        } else if (tid >= cw && tid < cw + 64 && tid < G && ((m_all >> (tid - cw)) & 1ull)) {
            const int c = tid, i = ROW[c];
            float *fc = F + static_cast<int64_t>(c) * ldf;
            const int8_t *sc = S + static_cast<int64_t>(c) * n;
            if (a.pv.kind == QA_KIND_SPARSE) {
                const float ds = DS[c];
                for (int q = L.adj_ptr[i]; q < L.adj_ptr[i + 1]; ++q)
                    fc[L.adj_idx[q]] += L.adj_val[q] * ds;
            } else {
                for (int q = L.inc_ptr[i]; q < L.inc_ptr[i + 1]; ++q) {
                    const int t = L.inc_idx[q];
                    float prod = L.t_coef[t];
                    for (int k = L.t_off[t]; k < L.t_off[t + 1]; ++k) prod *= sc[L.t_vars[k]];
                    for (int k = L.t_off[t]; k < L.t_off[t + 1]; ++k) {
                        const int v = L.t_vars[k];
                        if (v != i) fc[v] += 2.0f * prod * sc[v];
                    }
                }
            }
        }
    };
    // Apply the field updates of every flagged chain: one mask word per deciding wavefront.
    const int nwords = a.cmax >> 6;
    auto no_pre = []() {};
    auto apply_updates = [&]() {
        bool done = true;
        for (int w = 0; w < nwords; ++w) apply_word(flipmask[w], w << 6, false, no_pre, done);
    };
    // with skips, and pre() issued after the first batch's loads (or at the end when nothing flipped)
    auto apply_updates_pre = [&](auto &&pre) {
        bool done = false;
        for (int w = 0; w < nwords; ++w) apply_word(flipmask[w], w << 6, true, pre, done);
        if (!done) pre();
    };

    // Random permutation of 0..n-1 for this sweep (Fisher-Yates on one lane, Philox draws).
    auto make_order_into = [&](int *ord, uint32_t sweep_id) {
        if (tid == 0) shuffle_order(ord, n, cbase, sweep_id, a.k0, a.k1);
        __syncthreads();
    };
    auto make_order = [&](uint32_t sweep_id) { make_order_into(order, sweep_id); };

    // Row ring: the J rows of order positions [k, k + pfd] are staged in LDS slot (position % ring)
    // before they are needed; position k + pfd is fetched into registers during step k and stored
    // after that step's update, into the slot of a position no chain reads any more.
    constexpr int kPre = 8;  // registers per thread for one prefetched row (n <= 256 * kPre)
    float pre[kPre];
    auto ring_preload_from = [&](const int *order) {
        if (!a.ring) return;
        const int m = min(a.pfd, n);
        for (int q = 0; q < m; ++q) {
            const float *row = a.pv.J32 + static_cast<int64_t>(order[q]) * n;
            for (int j = tid; j < n; j += kThreads) Jring[static_cast<int64_t>(q % a.ring) * n + j] = row[j];
        }
        __syncthreads();
    };
    auto ring_preload = [&]() { ring_preload_from(order); };
    auto ring_fetch = [&](int pos) {
        if (!a.ring || pos >= n) return;
        const float *row = a.pv.J32 + static_cast<int64_t>(order[pos]) * n;
#pragma unroll
        for (int q = 0; q < kPre; ++q) {
            const int j = tid + q * kThreads;
            if (j < n) pre[q] = row[j];
        }
    };
    auto ring_store = [&](int pos) {
        if (!a.ring || pos >= n) return;
        float *dst = Jring + static_cast<int64_t>(pos % a.ring) * n;
#pragma unroll
        for (int q = 0; q < kPre; ++q) {
            const int j = tid + q * kThreads;
            if (j < n) dst[j] = pre[q];
        }
    };

    const int wave = tid >> 6, lane = tid & 63;

    // B of every replica of the group from the current spins; lanes tid < nrep accumulate their
    // replica's sample (count = false: only refresh blast, the value at the end of a step).
    double bacc = 0.0, b2acc = 0.0, blast = 0.0;
    auto bond_sample = [&](bool count) {
        for (int c = wave; c < G; c += kThreads / 64) {
            const int t = c % P;
            const int8_t *sc = S + static_cast<int64_t>(c) * n;
            const int8_t *sn = S + static_cast<int64_t>((c / P) * P + (t + 1 == P ? 0 : t + 1)) * n;
            int b = 0;
            for (int i = lane; i < n; i += 64) b += sc[i] * sn[i];
            for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o);
            if (lane == 0) ered[c] = static_cast<double>(b);
        }
        __syncthreads();
        if (tid < nrep) {
            double b = 0.0;
            for (int t = 0; t < P; ++t) b += ered[tid * P + t];
            if (count) {
                bacc += b;
                b2acc += b * b;
            }
            blast = b;
        }
        __syncthreads();
    };

    uint32_t sweep_id = a.sweep_base;
    for (int step = 0; step < a.steps; ++step) {
        const int pidx = a.param_of_rep ? (decider ? a.param_of_rep[r0 + rr_me] : 0) : step;
        const float bs = a.bs[pidx], jp = a.jp[pidx], join = a.join[pidx];
        // ---------------- slice sweeps (sqa_annealer.cpp:180-209) --------------------------
        // Skewed wavefront: at step tau, slice t works on order position k = tau - t. Slice t-1
        // has already passed position k (so s_{t-1}[i_k] is its value after this sweep) and slice
        // t+1 has not reached it (s_{t+1}[i_k] is still the value before this sweep) -- exactly
        // what the reference's slice-by-slice loop sees, including the ring closure where slice
        // P-1 reads slice 0's swept value. Simultaneous updates always touch distinct spin
        // indices, so they never interact. A sweep takes n + P - 1 steps.
        if (a.sweep_kind == 1) {
            // sqa_chi checkerboard (sqa_chi_annealer.cpp:81-140): each parity class of slices
            // sweeps sites 0..n-1 in order while the other class is held fixed; for odd P the
            // last slice (parity 0, neighbour of slice 0) forms its own phase after the other
            // even slices, which is the reference's single-thread cell order.
            for (int i = tid; i < n; i += kThreads) order[i] = i;
            __syncthreads();
            const int nph = P == 1 ? 1 : ((P & 1) ? 3 : 2);
            const int colour = P == 1 ? 0 : ((P & 1) ? (t_me == P - 1 ? 1 : ((t_me & 1) ? 2 : 0)) : (t_me & 1));
            for (int sw = 0; sw < a.spb; ++sw, ++sweep_id) {
                Philox4 rnd{0, 0, 0, 0};
                for (int ph = 0; ph < nph; ++ph) {
                    ring_preload();
                    for (int k = 0; k < n; ++k) {
                        ring_fetch(k + a.pfd);
                        const int tau = ph * n + k;
                        if (tid < a.cmax) {
                            const bool active = decider && colour == ph;
                            bool acc = false;
                            if ((tau & 3) == 0)
                                rnd = philox4x32_10(cbase + c_me, sweep_id, static_cast<uint32_t>(tau >> 2), TAG_CHECKER,
                                                    a.k0, a.k1);
                            if (active) {
                                int8_t *sc = S + static_cast<int64_t>(c_me) * n;
                                const int8_t s = sc[k];
                                const float f = F[static_cast<int64_t>(c_me) * ldf + k];
                                const int nb = S[static_cast<int64_t>(cprev) * n + k] + S[static_cast<int64_t>(cnext) * n + k];
                                const float fs = static_cast<float>(s);
                                const float de = fmaf(bs, -2.0f * fs * f, 2.0f * jp * fs * static_cast<float>(nb));
                                const uint32_t q = tau & 3;
                                const uint32_t word = q == 0 ? rnd.x : q == 1 ? rnd.y : q == 2 ? rnd.z : rnd.w;
                                acc = de <= 0.0f || u01(word) < __expf(-de);
                                DS[c_me] = -2.0f * fs;
                                ROW[c_me] = k;
                                if (a.ring) RSLOT[c_me] = k % a.ring;
                                if (acc) sc[k] = static_cast<int8_t>(-s);
                            }
                            const unsigned long long m = __ballot(acc);
                            if (lane == 0) flipmask[wave] = m;
                            nacc += acc ? 1ull : 0ull;
                        }
                        __syncthreads();
                        apply_updates();
                        ring_store(k + a.pfd);
                        __syncthreads();
                    }
                }
                if (a.mag_m2 && sw >= a.mag_burn) {
                    // m_r = sum of the replica's spins / (n P), accumulated as m^2 and |m| (:124-139)
                    for (int c = wave; c < G; c += kThreads / 64) {
                        const int8_t *sc = S + static_cast<int64_t>(c) * n;
                        int sum = 0;
                        for (int i = lane; i < n; i += 64) sum += sc[i];
                        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
                        if (lane == 0) ered[c] = static_cast<double>(sum);
                    }
                    __syncthreads();
                    if (tid < nrep) {
                        double tot = 0.0;
                        for (int t = 0; t < P; ++t) tot += ered[tid * P + t];
                        const double m = tot / (static_cast<double>(n) * P);
                        a.mag_m2[r0 + tid] += m * m;
                        a.mag_abs[r0 + tid] += fabs(m);
                    }
                    __syncthreads();
                }
            }
        } else
        for (int sw = 0; sw < a.spb; ++sw, ++sweep_id) {
            make_order(sweep_id);
            ring_preload();
            Philox4 rnd{0, 0, 0, 0};
            const int nsteps = n + P - 1;
            if constexpr (!kLds && NJ > 0) {
                // State in global memory: each chain's next decision inputs (field, own and
                // neighbour spins, and J[i_k][i_k+1] for its own flip) are loaded right after this
                // step's decisions, while the field updates run, so the two latencies overlap and a
                // decision reads registers only. The loads see every decision of this step (the
                // previous slice's spin at the next site included) and the updates of earlier steps;
                // this step's update of the chain's own row skips the loaded element, and the next
                // decision adds that flip itself -- the fmaf the update would apply -- and writes the
                // element back, so every field element gets the same fmaf sequence as in LDS.
                // With the spins in LDS (kSL) the decision reads them there, makes no global access
                // and leaves the element write-back to the update phase: there the chain's own row
                // update skips both the element just decided (written back with this step's flip,
                // J_ii = 0, applied by the same fmaf) and the element loaded for the next decision.
                float pf_f = 0.0f, pf_j = 0.0f, ds_prev = 0.0f;
                int pf_s = 0, pf_sp = 0, pf_sn = 0;
                bool fl_prev = false;
                int nx_i = 0, nx_i1 = -1;  // this step's site and the next one, for the loads
                bool wb = false;           // kSL: write wb_v to element wb_i of the chain's row
                int wb_i = 0;
                float wb_v = 0.0f;
                for (int tau = 0; tau < nsteps; ++tau) {
                    if (tid < a.cmax) {
                        const int k = tau - t_me;
                        const bool active = decider && k >= 0 && k < n;
                        bool acc = false;
                        nx_i1 = -1;
                        wb = false;
                        if ((tau & 3) == 0)
                            rnd = philox4x32_10(cbase + c_me, sweep_id, static_cast<uint32_t>(tau >> 2), TAG_SLICE,
                                                a.k0, a.k1);
                        if (active) {
                            const int i = order[k];
                            int8_t *sc = S + static_cast<int64_t>(c_me) * n;
                            int s, sp, sn;
                            float f;
                            if (k == 0) {
                                s = sc[i];
                                f = F[static_cast<int64_t>(c_me) * ldf + i];
                                sp = S[static_cast<int64_t>(cprev) * n + i];
                                sn = S[static_cast<int64_t>(cnext) * n + i];
                            } else if constexpr (kSL) {
                                s = sc[i];
                                f = fl_prev ? fmaf(ds_prev, pf_j, pf_f) : pf_f;
                                sp = S[static_cast<int64_t>(cprev) * n + i];
                                sn = S[static_cast<int64_t>(cnext) * n + i];
                            } else {
                                s = pf_s;
                                f = fl_prev ? fmaf(ds_prev, pf_j, pf_f) : pf_f;
                                if (fl_prev) F[static_cast<int64_t>(c_me) * ldf + i] = f;
                                sp = pf_sp;
                                sn = pf_sn;
                            }
                            const float fs = static_cast<float>(s);
                            const float de = fmaf(bs, -2.0f * fs * f, 2.0f * jp * fs * static_cast<float>(sp + sn));
                            const uint32_t q = tau & 3;
                            const uint32_t word = q == 0 ? rnd.x : q == 1 ? rnd.y : q == 2 ? rnd.z : rnd.w;
                            acc = de <= 0.0f || u01(word) < __expf(-de);
                            DS[c_me] = -2.0f * fs;
                            ROW[c_me] = i;
                            if (acc) sc[i] = static_cast<int8_t>(-s);
                            const bool fl_prev_in = fl_prev;
                            ds_prev = -2.0f * fs;
                            fl_prev = acc;
                            if (k + 1 < n) {
                                nx_i = i;
                                nx_i1 = order[k + 1];
                            }
                            SKIP[c_me] = nx_i1;
                            if constexpr (kSL) {
                                SKIP2[c_me] = i;
                                wb = fl_prev_in || acc;
                                wb_i = i;
                                wb_v = acc ? fmaf(-2.0f * fs, 0.0f, f) : f;
                            }
                        }
                        const unsigned long long m = __ballot(acc);
                        if (lane == 0) flipmask[wave] = m;
                        nacc += acc ? 1ull : 0ull;
                        // (spins in global memory) the field write-back and the spin store complete
                        // before the barrier: other wavefronts load those words in the update phase
                        if constexpr (!kSL) __builtin_amdgcn_s_waitcnt(0);
                    }
                    __syncthreads();
                    apply_updates_pre([&]() {
                        if (tid < a.cmax) {
                            if constexpr (kSL) {
                                if (wb) F[static_cast<int64_t>(c_me) * ldf + wb_i] = wb_v;
                            }
                            if (nx_i1 >= 0) {
                                pf_f = F[static_cast<int64_t>(c_me) * ldf + nx_i1];
                                pf_j = a.pv.J32[static_cast<int64_t>(nx_i) * n + nx_i1];
                                if constexpr (!kSL) {
                                    pf_s = S[static_cast<int64_t>(c_me) * n + nx_i1];
                                    pf_sp = S[static_cast<int64_t>(cprev) * n + nx_i1];
                                    pf_sn = S[static_cast<int64_t>(cnext) * n + nx_i1];
                                }
                            }
                        }
                    });
                    __syncthreads();
                }
                if (a.bstat) bond_sample(true);
                continue;
            }
            for (int tau = 0; tau < nsteps; ++tau) {
                ring_fetch(tau + a.pfd);
                if (tid < a.cmax) {
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
                        if (a.ring) RSLOT[c_me] = k % a.ring;
                        if (acc) sc[i] = static_cast<int8_t>(-s);
                    }
                    const unsigned long long m = __ballot(acc);
                    if (lane == 0) flipmask[wave] = m;
                    nacc += acc ? 1ull : 0ull;
                }
                __syncthreads();
                apply_updates();
                ring_store(tau + a.pfd);
                __syncthreads();
            }
            if (a.bstat) bond_sample(true);
        }
        // ---------------- cluster sweeps (sqa_annealer.cpp:215-288) ------------------------
        for (int sw = 0; sw < a.cl; ++sw, ++sweep_id) {
            make_order(sweep_id);
            ring_preload();
            for (int k = 0; k < n; ++k) {
                const int i = order[k];
                ring_fetch(k + a.pfd);
                if (decider && t_me == 0) {
                    // leader lane of replica rr_me grows the imaginary-time cluster (prev before
                    // next), membership as a bit mask over the replica's slices
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
                    if (P <= 64) {
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
                    } else {
                        // P > 64 (one replica per group): the same growth with multi-word masks in
                        // LDS, membership in cmask[0 .. nw), the stack in cmask[8 .. 8 + nw); the
                        // highest stacked slice is taken first, as with one word
                        unsigned long long *inc = cmask, *stk = cmask + 8;
                        const int nw = (P + 63) >> 6;
                        for (int w = 0; w < nw; ++w) inc[w] = stk[w] = 0ull;
                        auto has = [&](const unsigned long long *m, int t) { return (m[t >> 6] >> (t & 63)) & 1ull; };
                        auto put = [&](unsigned long long *m, int t) { m[t >> 6] |= 1ull << (t & 63); };
                        put(inc, seed);
                        put(stk, seed);
                        int top = seed >> 6;
                        for (;;) {
                            while (top >= 0 && stk[top] == 0ull) --top;
                            if (top < 0) break;
                            const int t = (top << 6) + 63 - __clzll(static_cast<long long>(stk[top]));
                            stk[top] &= ~(1ull << (t & 63));
                            const int pv = t == 0 ? P - 1 : t - 1, nx = t + 1 == P ? 0 : t + 1;
                            if (!has(inc, pv) && S[static_cast<int64_t>(base + pv) * n + i] == s0 &&
                                u01(next_u()) < join) { put(inc, pv); put(stk, pv); top = max(top, pv >> 6); }
                            if (!has(inc, nx) && S[static_cast<int64_t>(base + nx) * n + i] == s0 &&
                                u01(next_u()) < join) { put(inc, nx); put(stk, nx); top = max(top, nx >> 6); }
                        }
                        float dcl = 0.0f, dtm = 0.0f;
                        for (int t = 0; t < P; ++t) {
                            if (!has(inc, t)) continue;
                            dcl += -2.0f * s0 * F[static_cast<int64_t>(base + t) * ldf + i];
                            const int pv = t == 0 ? P - 1 : t - 1, nx = t + 1 == P ? 0 : t + 1;
                            if (!has(inc, pv)) dtm += 2.0f * jp * s0 * S[static_cast<int64_t>(base + pv) * n + i];
                            if (!has(inc, nx)) dtm += 2.0f * jp * s0 * S[static_cast<int64_t>(base + nx) * n + i];
                        }
                        const float de = fmaf(bs, dcl, dtm);
                        const bool ok = de <= 0.0f || u01(next_u()) < __expf(-de);
                        if (!ok)
                            for (int w = 0; w < nw; ++w) inc[w] = 0ull;
                    }
                }
                __syncthreads();
                bool acc = false;
                if (decider) {
                    acc = P <= 64 ? (cmask[rr_me] >> t_me) & 1ull : (cmask[t_me >> 6] >> (t_me & 63)) & 1ull;
                    int8_t *sc = S + static_cast<int64_t>(c_me) * n;
                    const int8_t s = sc[i];
                    DS[c_me] = -2.0f * s;
                    ROW[c_me] = i;
                    if (a.ring) RSLOT[c_me] = k % a.ring;
                    if (acc) sc[i] = static_cast<int8_t>(-s);
                }
                if (tid < a.cmax) {
                    const unsigned long long m = __ballot(acc);
                    if (lane == 0) flipmask[wave] = m;
                }
                __syncthreads();
                apply_updates();
                ring_store(k + a.pfd);
                __syncthreads();
            }
            if (a.bstat) bond_sample(true);
        }
        // ---------------- worldline sweeps (sqa_annealer.cpp:294-315) -----------------------
        for (int sw = 0; sw < a.wl; ++sw, ++sweep_id) {
            make_order(sweep_id);
            ring_preload();
            for (int k = 0; k < n; ++k) {
                const int i = order[k];
                ring_fetch(k + a.pfd);
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
                    if (a.ring) RSLOT[c_me] = k % a.ring;
                    if (acc) sc[i] = static_cast<int8_t>(-s);
                }
                if (tid < a.cmax) {
                    const unsigned long long m = __ballot(acc);
                    if (lane == 0) flipmask[wave] = m;
                }
                __syncthreads();
                apply_updates();
                ring_store(k + a.pfd);
                __syncthreads();
            }
        }
        // ---------------- bond statistics of the step ----------------------------------------
        if (a.bstat) {
            if (a.sweep_kind != 0 || a.spb + a.cl == 0) bond_sample(false);
            if (tid < 64) {  // wave 0: nrep <= 64 lanes hold the group's sums
                double x = bacc, y = b2acc;
                for (int o = 32; o > 0; o >>= 1) {
                    x += __shfl_xor(x, o);
                    y += __shfl_xor(y, o);
                }
                if (tid == 0) {
                    a.bstat[(static_cast<int64_t>(step) * a.ngroups + g) * 2] = x;
                    a.bstat[(static_cast<int64_t>(step) * a.ngroups + g) * 2 + 1] = y;
                    if (g == 0 && a.bend) a.bend[step] = blast;
                }
            }
            bacc = 0.0;
            b2acc = 0.0;
        }
        // ---------------- step tail ---------------------------------------------------------
        if (a.tail_mode != 2) group_tail(a, F, S, ldf, G, r0, g, step, ered, scal, gbest);
    }
    // final spins out (for the caller's state readback / diagnostics)
    if constexpr (kLds) group_save(a, F, S, ldf, G, r0);
    if constexpr (kSL) {
        __syncthreads();
        for (int e = tid; e < G * n; e += kThreads) a.gS[static_cast<int64_t>(r0) * P * n + e] = S[e];
    }
    if (a.cmax == kMaxChains) {
        if (tid < kMaxChains) {
            for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
            if (tid == 0) a.accepted[g] = nacc;
        }
    } else {
        // one partial per deciding wavefront (cmask is free after the last step), summed by thread 0
        if (tid < a.cmax) {
            for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
            if (lane == 0) cmask[wave] = nacc;
        }
        __syncthreads();
        if (tid == 0) {
            unsigned long long tot = 0;
            for (int w = 0; w < nwords; ++w) tot += cmask[w];
            a.accepted[g] = tot;
        }
    }
}


template <bool kLds, int NJ = 0, bool kSL = false>
__global__ void __launch_bounds__(kThreads) sqa_fp32_kernel(SqaArgs a) {
    general_group<kLds, NJ, kSL>(a, static_cast<int>(blockIdx.x));
}

// ---- look-ahead slice sweeps (dense, fields in global memory, spins in LDS) ------------------
// One replica group of plain slice sweeps (no worldline, cluster or checkerboard moves, no bond
// statistics) with one wavefront of chains: the same visit orders, random words and FP32
// arithmetic as the two-barrier kernel, so the same results bit for bit. NJ: 4-float groups per
// updating lane in a row (n <= 1024 NJ).
template <int NJ>
__device__ void la_group(const SqaArgs &a, const int g) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int n = a.pv.n, P = a.P;
    const int r0 = g * a.rg;
    const int nrep = min(a.rg, a.R - r0);
    const int G = nrep * P;
    const SqaLds L = carve_lds(a, smem, false);
    float *F = a.gF + static_cast<int64_t>(r0) * P * n;
    int8_t *S = L.S;
    int *order = L.order;
    const uint32_t cbase = static_cast<uint32_t>(r0 * P);
    group_init(a, F, S, a.ldf, G, r0, false);
    double gbest = (tid == 0 && a.tail_mode == 0) ? a.best_e[g] : 0.0;
    unsigned long long nacc = 0;
    const bool decider = tid < G;
    const int c_me = tid;
    const int rr_me = c_me / P, t_me = c_me - (c_me / P) * P;
    const int cprev = rr_me * P + (t_me == 0 ? P - 1 : t_me - 1);
    const int cnext = rr_me * P + (t_me + 1 == P ? 0 : t_me + 1);
    uint32_t sweep_id = a.sweep_base;
    for (int step = 0; step < a.steps; ++step) {
        const int pidx = a.param_of_rep ? (decider ? a.param_of_rep[r0 + rr_me] : 0) : step;
        const float bs = a.bs[pidx], jp = a.jp[pidx];
        for (int sw = 0; sw < a.spb; ++sw, ++sweep_id) {
            if (tid == 0) shuffle_order(order, n, cbase, sweep_id, a.k0, a.k1);
            __syncthreads();
            const int nsteps = n + P - 1;
            // Fields in global memory, spins in LDS, one wavefront of chains (chain c = lane c).
            // Roles, each running its own loop with two barriers per lockstep step tau:
            //   wave 0 (decisions): decides step tau from LDS only -- spins, the field element
            //     PF[c] and the couplings JC[.][c] between the site and the chain's last kLaCorr
            //     sites -- and publishes the flips (DS, ROW, flip mask) in the set tau & 1;
            //   waves 1-4 (updates, after the first barrier): finish the row updates of step
            //     tau - 1 (rows loaded during the previous update phase, the chains beyond those
            //     loaded and stored now), then load the rows of up to 8 / NJ flipped chains of
            //     step tau, which the next update phase finishes;
            //   wave 5 (helper, before the first barrier): stores into LDS, for decision tau + 1,
            //     what it loaded at step tau + 1 - kLaLead, then loads for decision tau + kLaLead
            //     the field element and the kLaCorr couplings (unconditional loads, clamped
            //     indexes), and evaluates the chains' Philox blocks.
            // Every flip of step s is stored during update phase s + 1, and no field element is
            // stored during a decision phase, so the element loaded at step tau holds exactly the
            // chain's flips up to step tau - 2. Decision tau + kLaLead adds its flips at steps
            // tau - 1 .. tau + kLaLead - 1 with the fmaf the row update applies, in the same order,
            // and so decides on the same bits as the element in memory (and as every other sweep
            // kernel). A row update of chain c loads F[c] after the store of c's previous flip by
            // the same lane. kLaLead leading steps make the first loads and trailing steps without
            // decisions let the last flips be stored; the step count is a multiple of kLaLead - 1
            // (the helper's register sets rotate with it).
            constexpr int kUpdT = 256;                          // threads of waves 1-4
            constexpr int kPreB = 8 / NJ;  // chains whose rows are loaded ahead (64 registers)
            constexpr int kNowB = NJ >= 8 ? 1 : 8 / NJ;  // chains per synchronous batch
            constexpr int kSets = kLaLead - 1;             // helper register sets in flight
            float *const LA = L.la;
            const int n4 = n >> 2;
            f4 *F4g = reinterpret_cast<f4 *>(F);
            const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
            const int tbeg = -kLaLead;
            const int tend = nsteps + 1 + (kSets - (nsteps + 1 - tbeg) % kSets) % kSets;
            // the first update phase reads the flip-mask set of step tbeg - 1: no flips
            if (tid < 2) reinterpret_cast<unsigned long long *>(LA + kLaFM)[tid] = 0ull;
            __syncthreads();
            __builtin_amdgcn_s_waitcnt(0);
            if (wv == 0) {
                // ---- decisions --------------------------------------------------------------
                // fl[j], dsh[j]: the chain's attempt at position k - 1 - j (flipped?, -2 s_old)
                bool fl[kLaCorr];
                float dsh[kLaCorr];
#pragma unroll
                for (int j = 0; j < kLaCorr; ++j) {
                    fl[j] = false;
                    dsh[j] = 0.0f;
                }
                const uint32_t *RWd = reinterpret_cast<const uint32_t *>(LA + kLaRW);
                int i_nx = order[0];  // the site of the next step, read one step ahead
                for (int tau = tbeg; tau < tend; ++tau) {
                    const int set = tau & 1;
                    const int k = tau - t_me;
                    const bool active = decider && k >= 0 && k < n;
                    bool acc = false;
                    const int i = i_nx;
                    i_nx = order[min(max(k + 1, 0), n - 1)];
                    // word tau & 3 of the chain's Philox block tau >> 2 (counter (chain, sweep,
                    // block, TAG_SLICE)), evaluated by the helper wave at least one step earlier
                    const uint32_t word = RWd[(((tau >> 2) & 1) * 4 + (tau & 3)) * 64 + c_me];
                    if (active) {
                        int8_t *sc = S + static_cast<int64_t>(c_me) * n;
                        // every LDS read of the decision is issued before the arithmetic starts
                        const int s = sc[i];
                        const int sp = S[static_cast<int64_t>(cprev) * n + i];
                        const int sn = S[static_cast<int64_t>(cnext) * n + i];
                        float f = LA[set * 64 + c_me];
                        float jc[kLaCorr];
#pragma unroll
                        for (int d = 0; d < kLaCorr; ++d) jc[d] = LA[kLaJC + (d * 2 + set) * 64 + c_me];
                        // JC slot d couples the site to position k - kLaCorr + d: oldest flip first
#pragma unroll
                        for (int d = 0; d < kLaCorr; ++d) {
                            const int j = kLaCorr - 1 - d;
                            f = fl[j] ? fmaf(dsh[j], jc[d], f) : f;
                        }
                        const float fs = static_cast<float>(s);
                        const float de = fmaf(bs, -2.0f * fs * f, 2.0f * jp * fs * static_cast<float>(sp + sn));
                        acc = de <= 0.0f || u01(word) < __expf(-de);
                        LA[kLaDS + set * 64 + c_me] = -2.0f * fs;
                        reinterpret_cast<int *>(LA + kLaROW)[set * 64 + c_me] = i;
                        if (acc) sc[i] = static_cast<int8_t>(-s);
#pragma unroll
                        for (int j = kLaCorr - 1; j > 0; --j) {
                            fl[j] = fl[j - 1];
                            dsh[j] = dsh[j - 1];
                        }
                        fl[0] = acc;
                        dsh[0] = -2.0f * fs;
                    }
                    const unsigned long long m = __ballot(acc);
                    if (lane == 0) reinterpret_cast<unsigned long long *>(LA + kLaFM)[set] = m;
                    nacc += acc ? 1ull : 0ull;
                    __syncthreads();
                    __syncthreads();
                }
            } else if (wv == kHelperWave) {
                // ---- helper: field elements and couplings kLaLead steps ahead, Philox words ---
                struct Ahead {
                    float f, j[kLaCorr];
                };
                const int t_l = lane % P;
                const bool ch = lane < G;
                const int cc = ch ? lane : 0;
                uint32_t *RWh = reinterpret_cast<uint32_t *>(LA + kLaRW);
                auto helper = [&](const int tau, Ahead &h) {
                    // Philox block b of chain = lane for steps 4b .. 4b + 3: block 0 at the first
                    // step, block b + 1 at tau = 4b, into the word set b & 1 the decisions of those
                    // steps read (the decisions of steps 4b .. 4b + 3 read the other set)
                    if (tau == tbeg || (tau >= 0 && (tau & 3) == 0)) {
                        const uint32_t b = tau < 0 ? 0u : static_cast<uint32_t>(tau >> 2) + 1u;
                        const Philox4 r4 = philox4x32_10(cbase + lane, sweep_id, b, TAG_SLICE, a.k0, a.k1);
                        uint32_t *dst = RWh + (b & 1u) * 256 + lane;
                        dst[0] = r4.x;
                        dst[64] = r4.y;
                        dst[128] = r4.z;
                        dst[192] = r4.w;
                    }
                    const int set = (tau + 1) & 1;
                    const int k1 = tau + 1 - t_l;
                    if (ch && k1 >= 0 && k1 < n) {
                        LA[set * 64 + lane] = h.f;
#pragma unroll
                        for (int d = 0; d < kLaCorr; ++d) LA[kLaJC + (d * 2 + set) * 64 + lane] = h.j[d];
                    }
                    const int kq = tau + kLaLead - t_l;
                    const int64_t col = order[min(max(kq, 0), n - 1)];
                    h.f = F[static_cast<int64_t>(cc) * n + col];
#pragma unroll
                    for (int d = 0; d < kLaCorr; ++d)
                        h.j[d] = a.pv.J32[static_cast<int64_t>(order[min(max(kq - kLaCorr + d, 0), n - 1)]) * n + col];
                    __syncthreads();
                    __syncthreads();
                };
                Ahead hs[kSets];
#pragma unroll
                for (int q = 0; q < kSets; ++q) {
                    hs[q].f = 0.0f;
#pragma unroll
                    for (int d = 0; d < kLaCorr; ++d) hs[q].j[d] = 0.0f;
                }
                for (int tau = tbeg; tau < tend; tau += kSets)
                    static_for<kSets>([&](auto q) { helper(tau + decltype(q)::value, hs[decltype(q)::value]); });
            } else {
                // ---- row updates ------------------------------------------------------------
                const int ut = tid - 64;  // 0 .. 255
                int nu_w = 0;             // 4-float groups per lane this wave holds (wave-uniform)
#pragma unroll
                for (int u = 0; u < NJ; ++u)
                    if ((wv - 1) * 64 + kUpdT * u < n4) nu_w = u + 1;
                nu_w = __builtin_amdgcn_readfirstlane(nu_w);
                // rows of the chains loaded ahead: J and F groups per chain and group index
                f4 prj[kPreB][NJ], prf[kPreB][NJ];
                int pcs[kPreB];
                int pcnt = 0;
                // synchronous update of B chains (the ones beyond those loaded ahead)
                auto upd_now = [&](auto bsz, auto nu, const int *cs, const int set, const int real) {
                    constexpr int B = decltype(bsz)::value;
                    constexpr int NU = decltype(nu)::value;
                    float dsv[B];
                    const f4 *rows[B];
                    f4 *frow[B];
#pragma unroll
                    for (int q = 0; q < B; ++q) {
                        dsv[q] = LA[kLaDS + set * 64 + cs[q]];
                        rows[q] = reinterpret_cast<const f4 *>(
                            a.pv.J32 + static_cast<int64_t>(reinterpret_cast<const int *>(LA + kLaROW)[set * 64 + cs[q]]) * n);
                        frow[q] = F4g + static_cast<int64_t>(cs[q]) * n4;
                    }
                    f4 rv[B][NU], fv[B][NU];
#pragma unroll
                    for (int u = 0; u < NU; ++u) {
                        const int gi = min(ut + kUpdT * u, n4 - 1);
#pragma unroll
                        for (int q = 0; q < B; ++q) {
                            rv[q][u] = rows[q][gi];
                            fv[q][u] = frow[q][gi];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < NU; ++u) {
                        const int gi = ut + kUpdT * u;
#pragma unroll
                        for (int q = 0; q < B; ++q) {
                            const f4 x = rv[q][u], y = fv[q][u];
                            f4 v = f4{fmaf(dsv[q], x.x, y.x), fmaf(dsv[q], x.y, y.y), fmaf(dsv[q], x.z, y.z),
                                      fmaf(dsv[q], x.w, y.w)};
                            // computed on every path (not only under the store's mask): the
                            // loads are then waited for here, never later
                            asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
                            if (q < real && gi < n4) frow[q][gi] = v;
                        }
                    }
                };
                auto upd_now_nu = [&](auto bsz, const int *cs, const int set, const int real) {
                    static_for<NJ>([&](auto u) {
                        if (nu_w == decltype(u)::value + 1)
                            upd_now(bsz, std::integral_constant<int, decltype(u)::value + 1>{}, cs, set, real);
                    });
                };
                // finish the chains loaded ahead (their flips are in set `set`)
                auto finish_pre = [&](const int set) {
                    // every register a load of the previous phase may have written is read here
                    // on every path, so the wait for those loads sits here and nowhere later
                    static_for<kPreB>([&](auto qq) {
#pragma unroll
                        for (int u = 0; u < NJ; ++u) {
                            const f4 x = prj[decltype(qq)::value][u], y = prf[decltype(qq)::value][u];
                            asm volatile("" ::"v"(x.x), "v"(x.y), "v"(x.z), "v"(x.w), "v"(y.x), "v"(y.y), "v"(y.z),
                                         "v"(y.w));
                        }
                    });
                    static_for<kPreB>([&](auto qq) {
                        constexpr int q = decltype(qq)::value;
                        if (q < pcnt) {
                            const float ds = LA[kLaDS + set * 64 + pcs[q]];
                            f4 *frow = F4g + static_cast<int64_t>(pcs[q]) * n4;
#pragma unroll
                            for (int u = 0; u < NJ; ++u) {
                                if (u < nu_w) {
                                    const int gi = ut + kUpdT * u;
                                    const f4 x = prj[q][u], y = prf[q][u];
                                    f4 v = f4{fmaf(ds, x.x, y.x), fmaf(ds, x.y, y.y), fmaf(ds, x.z, y.z),
                                              fmaf(ds, x.w, y.w)};
                                    asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
                                    if (gi < n4) frow[gi] = v;
                                }
                            }
                        }
                    });
                };
                // load the rows of the first flipped chains of set `set`
                auto load_pre = [&](uint64_t &m, const int set) {
                    pcnt = 0;
                    static_for<kPreB>([&](auto qq) {
                        constexpr int q = decltype(qq)::value;
                        if (m) {
                            const int c = __builtin_ctzll(m);
                            m &= m - 1u;
                            pcs[q] = c;
                            pcnt = q + 1;
                            const f4 *row = reinterpret_cast<const f4 *>(
                                a.pv.J32 + static_cast<int64_t>(reinterpret_cast<const int *>(LA + kLaROW)[set * 64 + c]) * n);
                            const f4 *frow = F4g + static_cast<int64_t>(c) * n4;
#pragma unroll
                            for (int u = 0; u < NJ; ++u) {
                                if (u < nu_w) {
                                    const int gi = min(ut + kUpdT * u, n4 - 1);
                                    prj[q][u] = row[gi];
                                    prf[q][u] = frow[gi];
                                }
                            }
                        }
                    });
                };
                auto mask_of = [&](const int set) -> uint64_t {
                    const unsigned long long m_all = reinterpret_cast<const unsigned long long *>(LA + kLaFM)[set];
                    const uint32_t mlo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all));
                    const uint32_t mhi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all >> 32));
                    return (static_cast<uint64_t>(mhi) << 32) | mlo;
                };
                for (int tau = tbeg; tau < tend; ++tau) {
                    const int set = tau & 1;
                    __syncthreads();
                    if (nu_w > 0) {
                        // step tau - 1: the chains loaded ahead, then the rest of its flips
                        finish_pre(set ^ 1);
                        uint64_t mp = mask_of(set ^ 1);
                        for (int q = 0; q < pcnt; ++q) mp &= mp - 1u;
                        while (mp) {
                            // batches of 4 while at least 3 chains remain: one round trip per 4
                            // chains; a slot beyond the remaining chains loads chain cs[0]'s rows
                            // again and stores nothing
                            const int cnt = __popcll(mp);
                            int cs[4] = {0, 0, 0, 0};
                            const int b = (kNowB >= 4 && cnt >= 3) ? 4 : (kNowB >= 2 && cnt >= 2) ? 2 : 1;
                            const int real = min(b, cnt);
                            static_for<4>([&](auto qq) {
                                constexpr int q = decltype(qq)::value;
                                if (q < real) {
                                    cs[q] = __builtin_ctzll(mp);
                                    mp &= mp - 1u;
                                } else if (q < b) {
                                    cs[q] = cs[0];
                                }
                            });
                            if (kNowB >= 4 && b == 4) upd_now_nu(std::integral_constant<int, (kNowB >= 4 ? 4 : 1)>{}, cs, set ^ 1, real);
                            else if (kNowB >= 2 && b == 2) upd_now_nu(std::integral_constant<int, (kNowB >= 2 ? 2 : 1)>{}, cs, set ^ 1, real);
                            else upd_now_nu(std::integral_constant<int, 1>{}, cs, set ^ 1, real);
                        }
                        // step tau: load ahead
                        uint64_t mc = mask_of(set);
                        load_pre(mc, set);
                    }
                    __syncthreads();
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        if (a.tail_mode != 2) group_tail(a, F, S, a.ldf, G, r0, g, step, L.ered, L.scal, gbest);
    }
    __syncthreads();
    for (int e = tid; e < G * n; e += kThreads) a.gS[static_cast<int64_t>(r0) * P * n + e] = S[e];
    if (tid < kMaxChains) {
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        if (tid == 0) a.accepted[g] = nacc;
    }
}

template <int NJ>
__global__ void __launch_bounds__(kThreads) sqa_la_kernel(SqaArgs a) {
    la_group<NJ>(a, static_cast<int>(blockIdx.x));
}

// ---- pipelined slice sweeps (dense couplings, LDS-resident state) ---------------------------
// The same visit orders, random words and FP32 arithmetic as the two-barrier slice sweep of
// sqa_fp32_kernel, hence bit-identical results, with one barrier per lockstep step and the
// workgroup's waves in fixed roles:
//   wave 0     decides step tau: lane c = chain c (sqa_annealer.cpp:188-208 acceptance);
//   waves 1-4  apply the field updates f_c += J[i,:] ds_c of step tau-1, the flipped chains
//              dealt round-robin to the four waves, 16-byte LDS accesses;
//   wave 5     stages the J row of order position tau + pfd into the LDS ring (loaded during
//              step tau, stored during step tau + 1) and takes one Fisher-Yates step of the next
//              sweep's visit order, whose draw blocks it evaluates 64 at a time.
// The decision of chain c at site i = order[k] needs f_c[i] including c's own flip at step
// tau-1 (site order[k-1]). The updaters skip the 4-float group holding i in that update and
// wave 0 applies the same fmaf to the group itself, reading J[order[k-1]] from the ring, so no
// LDS word is written by two waves in one step. NQ: 4-float groups per lane in a row (ldf <= 256 NQ).
template <int NQ>
__device__ void pipe_group(const SqaArgs &a, const int g) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int kB = NQ >= 8 ? 1 : 8 / NQ;  // flipped chains per update batch
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: role branches stay scalar
    const int n = a.pv.n, P = a.P;
    const int r0 = g * a.rg;
    const int nrep = min(a.rg, a.R - r0);
    const int G = nrep * P;
    const int ldf = a.ldf;  // == a.ldr, a multiple of 4
    const int ngr = ldf >> 2;
    const int W = a.ring;
    const SqaLds L = carve_lds(a, smem, true);
    // Philox words of the current and the next four-step block, [block parity][word][chain]; the
    // two-barrier sweeps' per-chain scratch (RSLOT .. rflag, 2304 contiguous bytes) is unused here
    uint32_t *RW = reinterpret_cast<uint32_t *>(L.RSLOT);
    float *F = L.F;
    int8_t *S = L.S;
    f4 *F4 = reinterpret_cast<f4 *>(L.F);
    const f4 *J4 = reinterpret_cast<const f4 *>(L.Jring);
    f4 *Jring4 = reinterpret_cast<f4 *>(L.Jring);
    unsigned long long *flipmask = L.flipmask;
    const uint32_t cbase = static_cast<uint32_t>(r0 * P);

    group_init(a, F, S, ldf, G, r0, true);
    double gbest = (tid == 0 && a.tail_mode == 0) ? a.best_e[g] : 0.0;
    unsigned long long nacc = 0;

    const bool decider = tid < G;
    const int c_me = tid;
    const int rr_me = c_me / P, t_me = c_me - (c_me / P) * P;
    const int cprev = rr_me * P + (t_me == 0 ? P - 1 : t_me - 1);
    const int cnext = rr_me * P + (t_me + 1 == P ? 0 : t_me + 1);
    const int t_lane = lane % P;

    // wave 5: one staged row, NQ groups per lane
    f4 hpre[NQ];
    // (every lane loads NQ groups, clamped into the row, so that hpre stays in registers)
    auto ring_fetch = [&](const int *ord, int pos) {
        const f4 *row = reinterpret_cast<const f4 *>(a.pv.J32p + static_cast<int64_t>(ord[min(pos, n - 1)]) * ldf);
        static_for<NQ>([&](auto r) { hpre[r] = row[min(lane + 64 * r, ngr - 1)]; });
    };
    auto ring_store = [&](int pos, int slot) {
        if (pos >= n) return;
        static_for<NQ>([&](auto r) {
            const int gi = lane + 64 * r;
            if (gi < ngr) Jring4[slot * ngr + gi] = hpre[r];
        });
    };

    // waves 1-4: the updates of step tau-1 (flip masks of parity bb). Chain c was at order
    // position kc = tau - 1 - t_c (ring slot slot_l on lane c) and decides site order[kc + 1]
    // next; its group is left to wave 0.
    f4 *junk4 = reinterpret_cast<f4 *>(L.junk) + lane;
    auto apply = [&](int bb, int slot_l, int skip_l) {
        const unsigned long long m_all = flipmask[bb];
        const unsigned long long u_all = flipmask[2 + bb];
        const uint32_t mlo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all));
        const uint32_t mhi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all >> 32));
        const uint32_t ulo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u_all));
        const uint32_t uhi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u_all >> 32));
        const unsigned long long up = (static_cast<unsigned long long>(uhi) << 32) | ulo;
        // this wave's share: the flipped chains c with c % 4 == wave - 1
        unsigned long long mine = ((static_cast<unsigned long long>(mhi) << 32) | mlo) & (0x1111111111111111ull << (wave - 1));
        auto batch = [&](auto bsz) {
            constexpr int B = decltype(bsz)::value;
            float dsv[B];
            int roff[B], foff[B], skp[B];
#pragma unroll
            for (int q = 0; q < B; ++q) {
                const int c = __ffsll(static_cast<long long>(mine)) - 1;
                mine &= mine - 1ull;
                dsv[q] = ((up >> c) & 1ull) ? -2.0f : 2.0f;
                roff[q] = __builtin_amdgcn_readlane(slot_l, c) * ngr;
                skp[q] = __builtin_amdgcn_readlane(skip_l, c);
                foff[q] = c * ngr;
            }
            // every load and store is unconditional (loads clamped into the row, excluded stores
            // sent to junk), so a batch is one LDS round trip with no branches
            f4 rv[B][NQ], fv[B][NQ];
            static_for<B>([&](auto q) {
                static_for<NQ>([&](auto r) {
                    const int gi = min(lane + 64 * r, ngr - 1);
                    rv[q][r] = J4[roff[q] + gi];
                    fv[q][r] = F4[foff[q] + gi];
                });
            });
            static_for<B>([&](auto q) {
                static_for<NQ>([&](auto r) {
                    const int gi = lane + 64 * r;
                    const f4 x = rv[q][r], y = fv[q][r];
                    f4 *dst = (gi < ngr && gi != skp[q]) ? F4 + foff[q] + gi : junk4;
                    *dst = f4{fmaf(dsv[q], x.x, y.x), fmaf(dsv[q], x.y, y.y), fmaf(dsv[q], x.z, y.z),
                              fmaf(dsv[q], x.w, y.w)};
                });
            });
        };
        while (mine) {
            const int cnt = __popcll(mine);
            if (kB >= 8 && cnt >= 8) batch(std::integral_constant<int, (kB >= 8 ? 8 : 1)>{});
            else if (kB >= 4 && cnt >= 4) batch(std::integral_constant<int, (kB >= 4 ? 4 : 1)>{});
            else if (kB >= 2 && cnt >= 2) batch(std::integral_constant<int, (kB >= 2 ? 2 : 1)>{});
            else batch(std::integral_constant<int, 1>{});
        }
    };

    uint32_t sweep_id = a.sweep_base;
    bool have_next = false;  // order buffer cur already holds the order of sweep next_id
    uint32_t next_id = 0;
    int cur = 0;
    for (int step = 0; step < a.steps; ++step) {
        const int pidx = a.param_of_rep ? (decider ? a.param_of_rep[r0 + rr_me] : 0) : step;
        const float bs = a.bs[pidx], jp = a.jp[pidx];
        for (int sw = 0; sw < a.spb; ++sw, ++sweep_id) {
            int *ocur = L.order + cur * L.obstride;
            int *onext = L.order + (cur ^ 1) * L.obstride;
            if (!(have_next && next_id == sweep_id)) {
                if (tid == 0) shuffle_order(ocur, n, cbase, sweep_id, a.k0, a.k1);
                __syncthreads();
            }
            const bool pre_next = sw + 1 < a.spb || step + 1 < a.steps;
            if (pre_next)
                for (int i = tid; i < n; i += kThreads) onext[i] = i;
            if (wave == kHelperWave) {  // Philox block 0 of chain = lane
                const Philox4 r = philox4x32_10(cbase + lane, sweep_id, 0u, TAG_SLICE, a.k0, a.k1);
                RW[lane] = r.x;
                RW[64 + lane] = r.y;
                RW[128 + lane] = r.z;
                RW[192 + lane] = r.w;
            }
            // ring rows of positions 0 .. pfd-1
            for (int e = tid; e < min(a.pfd, n) * ngr; e += kThreads) {
                const int q = e / ngr, gi = e - q * ngr;
                Jring4[q * ngr + gi] = reinterpret_cast<const f4 *>(a.pv.J32p + static_cast<int64_t>(ocur[q]) * ldf)[gi];
            }
            __syncthreads();

            const int nsteps = n + P - 1;
            bool fl_prev = false;        // wave 0: chain flipped at the previous step
            float ds_prev = 0.0f;        // wave 0: -2 s_old of that flip
            int i_cur = ocur[0];         // wave 0: site of this step (once active)
            int sl = W - 1 - t_lane;     // waves 0-4: ring slot of position tau - 1 - t of lane's chain
            int skn = -1;                // waves 1-4: skip group of lane's chain in the next step
            PhiloxRun nxt{};                   // wave 5: block of the next four steps for chain = lane
            int sh_ii = pre_next ? n - 1 : 0;  // wave 5: next Fisher-Yates iteration
            Philox4 shb{0, 0, 0, 0};            // wave 5: this lane's draw block of the batch
            int hslot = a.pfd % W;              // wave 5: slot of position tau - 1 + pfd
            for (int tau = 0; tau <= nsteps; ++tau) {
                const int b = tau & 1;
                if (wave == 0) {
                    if (tau < nsteps) {
                        // every lane issues all of its loads at once (addresses kept valid: lanes
                        // without a chain read chain 0's), then decides with selects
                        const int k = tau - t_me;
                        const bool active = decider && k >= 0 && k < n;
                        const int cc = decider ? c_me : 0;
                        const int ccp = decider ? cprev : 0, ccn = decider ? cnext : 0;
                        const int i = i_cur;
                        const int gq = cc * ngr + (i >> 2);
                        const f4 fv = F4[gq];
                        const f4 jv = J4[sl * ngr + (i >> 2)];
                        const int8_t s = S[cc * n + i];
                        const int nb = S[ccp * n + i] + S[ccn * n + i];
                        const uint32_t word = RW[((tau >> 2) & 1) * 256 + (tau & 3) * 64 + cc];
                        const int i_nx = ocur[min(max(k + 1, 0), n - 1)];
                        const f4 fc = f4{fmaf(ds_prev, jv.x, fv.x), fmaf(ds_prev, jv.y, fv.y), fmaf(ds_prev, jv.z, fv.z),
                                         fmaf(ds_prev, jv.w, fv.w)};
                        const f4 fu = fl_prev ? fc : fv;
                        const int e = i & 3;
                        const float f = e == 0 ? fu.x : e == 1 ? fu.y : e == 2 ? fu.z : fu.w;
                        const float fs = static_cast<float>(s);
                        const float de = fmaf(bs, -2.0f * fs * f, 2.0f * jp * fs * static_cast<float>(nb));
                        const bool acc = active && (de <= 0.0f || u01(word) < __expf(-de));
                        if (active && fl_prev) F4[gq] = fc;
                        if (acc) S[cc * n + i] = static_cast<int8_t>(-s);
                        if (k + 1 >= 0 && k + 1 < n) i_cur = i_nx;
                        if (active) ds_prev = -2.0f * fs;
                        fl_prev = acc;
                        sl = sl + 1 == W ? 0 : sl + 1;
                        const unsigned long long m = __ballot(acc);
                        const unsigned long long mu = __ballot(acc && s > 0);
                        if (tid == 0) {
                            flipmask[b] = m;
                            flipmask[2 + b] = mu;
                        }
                        if (tid == 0) nacc += __popcll(m);
                    }
                } else if (wave <= 4) {
                    // lane's chain was at position tau - 1 - t in step tau - 1 and decides the site
                    // at position tau - t next: its 4-float group is wave 0's in this step
                    const int skip_l = skn;
                    {
                        const int kn = tau + 1 - t_lane;  // position after next, for the next step
                        skn = (lane < G && kn >= 1 && kn < n) ? ocur[kn] >> 2 : -1;
                    }
                    if (tau > 0) apply(b ^ 1, sl, skip_l);
                    sl = sl + 1 == W ? 0 : sl + 1;
                } else {
                    // Philox block m + 1 of chain = lane (counter (chain, sweep, m + 1, TAG_SLICE)),
                    // evaluated 3+3+2+2 rounds over the steps of block m and published at its last
                    {
                        const int q = tau & 3;
                        if (q == 0) {
                            nxt.start(cbase + lane, sweep_id, static_cast<uint32_t>((tau >> 2) + 1), TAG_SLICE, a.k0, a.k1);
                            nxt.rounds<3>();
                        } else if (q == 1) {
                            nxt.rounds<3>();
                        } else if (q == 2) {
                            nxt.rounds<2>();
                        } else {
                            nxt.rounds<2>();
                            uint32_t *dst = RW + (((tau >> 2) + 1) & 1) * 256 + lane;
                            dst[0] = nxt.c0;
                            dst[64] = nxt.c1;
                            dst[128] = nxt.c2;
                            dst[192] = nxt.c3;
                        }
                    }
                    // store the row fetched last step, shuffle, fetch this step's row: the loads stay
                    // in flight for a whole step
                    if (tau > 0) {
                        ring_store(tau - 1 + a.pfd, hslot);
                        hslot = hslot + 1 == W ? 0 : hslot + 1;
                    }
                    if (sh_ii > 0) {
                        // iteration sh_ii of shuffle_order(onext, ..., sweep_id + 1, ...)
                        const int it = n - 1 - sh_ii;
                        if ((it & 255) == 0) {
                            const int i0 = n - 1 - 4 * ((it >> 2) + lane);
                            if (i0 > 0)
                                shb = philox4x32_10(cbase, sweep_id + 1, static_cast<uint32_t>(i0), TAG_ORDER, a.k0, a.k1);
                        }
                        const int qq = it & 3;
                        const uint32_t wv = qq == 0 ? shb.x : qq == 1 ? shb.y : qq == 2 ? shb.z : shb.w;
                        const uint32_t d = __builtin_amdgcn_readlane(wv, (it >> 2) & 63);
                        if (lane == 0) {
                            const int j = static_cast<int>(below(d, static_cast<uint32_t>(sh_ii + 1)));
                            const int tmp = onext[sh_ii];
                            onext[sh_ii] = onext[j];
                            onext[j] = tmp;
                        }
                        --sh_ii;
                    }
                    ring_fetch(ocur, tau + a.pfd);
                }
                __syncthreads();
            }
            have_next = pre_next;
            if (pre_next) {
                next_id = sweep_id + 1;
                cur ^= 1;
            }
        }
        if (a.tail_mode != 2) group_tail(a, F, S, ldf, G, r0, g, step, L.ered, L.scal, gbest);
    }
    group_save(a, F, S, ldf, G, r0);
    if (tid < kMaxChains) {
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        if (tid == 0) a.accepted[g] = nacc;
    }
}


template <int NQ>
__global__ void __launch_bounds__(kThreads) sqa_pipe_kernel(SqaArgs a) {
    pipe_group<NQ>(a, static_cast<int>(blockIdx.x));
}

// ---- pipelined worldline sweeps (dense couplings, LDS-resident state) ------------------------
// The worldline move of the two-barrier kernel (sqa_annealer.cpp:294-315: every slice of a replica
// flips site i together, accepted on beta/P times the summed classical change) with one barrier
// per site instead of four, the same visit orders, random words and FP32 arithmetic, hence
// bit-identical results. Roles:
//   wave 0     decides site k for every replica of the group: lane c = chain c computes
//              -2 s f of its slice, then every lane sums its replica's slices from LDS in slice
//              order (the two-barrier kernel's order), so the replica's lanes reach one decision;
//   waves 1-4  apply the field updates of site k-1 (one J row for every flipped chain, loaded
//              once per lane), the flipped chains dealt round-robin to the four waves;
//   wave 5     stages the J row of order position k + pfd into the LDS ring, evaluates the
//              decisions' Philox words a block of sites ahead and takes one Fisher-Yates step of
//              the next sweep's visit order.
// The decision at site k needs f_c[i_k] including the flip of site k-1: the updaters skip the
// 4-float group holding i_k (the same group for every chain) and wave 0 applies that flip's fmaf
// to the group itself, reading J[i_{k-1}] from the ring, so no LDS word is written by two waves in
// one step. Runs a.wl worldline sweeps per step (no slice or cluster sweeps), then the step tail.
template <int NQ>
__device__ void pipe_wl_group(const SqaArgs &a, const int g) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int kB = NQ >= 8 ? 1 : 8 / NQ;  // flipped chains per update batch
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n = a.pv.n, P = a.P;
    const int r0 = g * a.rg;
    const int nrep = min(a.rg, a.R - r0);
    const int G = nrep * P;
    const int ldf = a.ldf;  // == a.ldr, a multiple of 4
    const int ngr = ldf >> 2;
    const int W = a.ring;
    const int kb = 64 / a.rg;  // sites per block of decision words (>= 4)
    const SqaLds L = carve_lds(a, smem, true);
    // decision words of the current and the next block, [block parity][replica * kb + site]
    uint32_t *RW = reinterpret_cast<uint32_t *>(L.RSLOT);
    float *wred = L.wred;
    float *F = L.F;
    int8_t *S = L.S;
    f4 *F4 = reinterpret_cast<f4 *>(L.F);
    const f4 *J4 = reinterpret_cast<const f4 *>(L.Jring);
    f4 *Jring4 = reinterpret_cast<f4 *>(L.Jring);
    unsigned long long *flipmask = L.flipmask;
    const uint32_t cbase = static_cast<uint32_t>(r0 * P);

    group_init(a, F, S, ldf, G, r0, true);
    double gbest = (tid == 0 && a.tail_mode == 0) ? a.best_e[g] : 0.0;

    const bool decider = tid < G;
    const int c_me = decider ? tid : 0;
    const int rr_me = c_me / P;

    // wave 5: one staged row, NQ groups per lane (clamped into the row, so hpre stays in registers)
    f4 hpre[NQ];
    auto ring_fetch = [&](const int *ord, int pos) {
        const f4 *row = reinterpret_cast<const f4 *>(a.pv.J32p + static_cast<int64_t>(ord[min(pos, n - 1)]) * ldf);
        static_for<NQ>([&](auto r) { hpre[r] = row[min(lane + 64 * r, ngr - 1)]; });
    };
    auto ring_store = [&](int pos, int slot) {
        if (pos >= n) return;
        static_for<NQ>([&](auto r) {
            const int gi = lane + 64 * r;
            if (gi < ngr) Jring4[slot * ngr + gi] = hpre[r];
        });
    };

    // waves 1-4: the updates of site k-1 (flip masks of parity bb, ring slot `slot`); group gskip
    // (that of site k, or -1) is left to wave 0
    f4 *junk4 = reinterpret_cast<f4 *>(L.junk) + lane;
    auto apply = [&](int bb, int slot, int gskip) {
        const unsigned long long m_all = flipmask[bb];
        const unsigned long long u_all = flipmask[2 + bb];
        const uint32_t mlo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all));
        const uint32_t mhi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(m_all >> 32));
        const uint32_t ulo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u_all));
        const uint32_t uhi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u_all >> 32));
        const unsigned long long up = (static_cast<unsigned long long>(uhi) << 32) | ulo;
        unsigned long long mine = ((static_cast<unsigned long long>(mhi) << 32) | mlo) & (0x1111111111111111ull << (wave - 1));
        if (!mine) return;
        f4 rv[NQ];
        static_for<NQ>([&](auto r) { rv[r] = J4[slot * ngr + min(lane + 64 * r, ngr - 1)]; });
        auto batch = [&](auto bsz) {
            constexpr int B = decltype(bsz)::value;
            float dsv[B];
            int foff[B];
#pragma unroll
            for (int q = 0; q < B; ++q) {
                const int c = __ffsll(static_cast<long long>(mine)) - 1;
                mine &= mine - 1ull;
                dsv[q] = ((up >> c) & 1ull) ? -2.0f : 2.0f;
                foff[q] = c * ngr;
            }
            f4 fv[B][NQ];
            static_for<B>([&](auto q) {
                static_for<NQ>([&](auto r) { fv[q][r] = F4[foff[q] + min(lane + 64 * r, ngr - 1)]; });
            });
            static_for<B>([&](auto q) {
                static_for<NQ>([&](auto r) {
                    const int gi = lane + 64 * r;
                    const f4 x = rv[r], y = fv[q][r];
                    f4 *dst = (gi < ngr && gi != gskip) ? F4 + foff[q] + gi : junk4;
                    *dst = f4{fmaf(dsv[q], x.x, y.x), fmaf(dsv[q], x.y, y.y), fmaf(dsv[q], x.z, y.z),
                              fmaf(dsv[q], x.w, y.w)};
                });
            });
        };
        while (mine) {
            const int cnt = __popcll(mine);
            if (kB >= 8 && cnt >= 8) batch(std::integral_constant<int, (kB >= 8 ? 8 : 1)>{});
            else if (kB >= 4 && cnt >= 4) batch(std::integral_constant<int, (kB >= 4 ? 4 : 1)>{});
            else if (kB >= 2 && cnt >= 2) batch(std::integral_constant<int, (kB >= 2 ? 2 : 1)>{});
            else batch(std::integral_constant<int, 1>{});
        }
    };

    uint32_t sweep_id = a.sweep_base;
    bool have_next = false;  // order buffer cur already holds the order of sweep next_id
    uint32_t next_id = 0;
    int cur = 0;
    for (int step = 0; step < a.steps; ++step) {
        const int pidx = a.param_of_rep ? (decider ? a.param_of_rep[r0 + rr_me] : 0) : step;
        const float bs = a.bs[pidx];
        for (int sw = 0; sw < a.wl; ++sw, ++sweep_id) {
            int *ocur = L.order + cur * L.obstride;
            int *onext = L.order + (cur ^ 1) * L.obstride;
            if (!(have_next && next_id == sweep_id)) {
                if (tid == 0) shuffle_order(ocur, n, cbase, sweep_id, a.k0, a.k1);
                __syncthreads();
            }
            const bool pre_next = sw + 1 < a.wl || step + 1 < a.steps;
            if (pre_next)
                for (int i = tid; i < n; i += kThreads) onext[i] = i;
            if (wave == kHelperWave) {  // decision words of block 0: replica lane / kb, site lane % kb
                const int rr = lane / kb, q = lane - (lane / kb) * kb;
                RW[lane] = philox4x32_10(cbase + rr * P, sweep_id, static_cast<uint32_t>(q), TAG_WORLD, a.k0, a.k1).x;
            }
            // ring rows of positions 0 .. pfd-1
            for (int e = tid; e < min(a.pfd, n) * ngr; e += kThreads) {
                const int q = e / ngr, gi = e - q * ngr;
                Jring4[q * ngr + gi] = reinterpret_cast<const f4 *>(a.pv.J32p + static_cast<int64_t>(ocur[q]) * ldf)[gi];
            }
            __syncthreads();

            bool fl_prev = false;        // wave 0: the lane's replica flipped the previous site
            float ds_prev = 0.0f;        // wave 0: -2 s_old of that flip
            int i_cur = ocur[0];         // wave 0: site of this step
            int sl = W - 1;              // waves 0-4: ring slot of position tau - 1
            PhiloxRun nxt{};                   // wave 5: decision word of the next block for lane
            int sh_ii = pre_next ? n - 1 : 0;  // wave 5: next Fisher-Yates iteration
            Philox4 shb{0, 0, 0, 0};            // wave 5: this lane's draw block of the batch
            int hslot = a.pfd % W;              // wave 5: slot of position tau - 1 + pfd
            for (int tau = 0; tau <= n; ++tau) {
                const int b = tau & 1;
                if (wave == 0) {
                    if (tau < n) {
                        const int i = i_cur;
                        const int gq = c_me * ngr + (i >> 2);
                        const f4 fv = F4[gq];
                        const f4 jv = J4[sl * ngr + (i >> 2)];
                        const int8_t s = S[c_me * n + i];
                        const int kq = tau / kb;
                        const uint32_t word = RW[(kq & 1) * 64 + rr_me * kb + (tau - kq * kb)];
                        const int i_nx = ocur[min(tau + 1, n - 1)];
                        const f4 fc = f4{fmaf(ds_prev, jv.x, fv.x), fmaf(ds_prev, jv.y, fv.y), fmaf(ds_prev, jv.z, fv.z),
                                         fmaf(ds_prev, jv.w, fv.w)};
                        const f4 fu = fl_prev ? fc : fv;
                        const int e = i & 3;
                        const float f = e == 0 ? fu.x : e == 1 ? fu.y : e == 2 ? fu.z : fu.w;
                        if (decider) wred[c_me] = -2.0f * static_cast<float>(s) * f;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        float dcl = 0.0f;  // the replica's slices in slice order
                        const float *wr = wred + rr_me * P;
#pragma unroll 8
                        for (int t = 0; t < P; ++t) dcl += wr[t];
                        const float de = bs * dcl;
                        const bool acc = decider && (de <= 0.0f || u01(word) < __expf(-de));
                        if (decider && fl_prev) F4[gq] = fc;
                        if (acc) S[c_me * n + i] = static_cast<int8_t>(-s);
                        i_cur = i_nx;
                        ds_prev = -2.0f * static_cast<float>(s);
                        fl_prev = acc;
                        sl = sl + 1 == W ? 0 : sl + 1;
                        const unsigned long long m = __ballot(acc);
                        const unsigned long long mu = __ballot(acc && s > 0);
                        if (tid == 0) {
                            flipmask[b] = m;
                            flipmask[2 + b] = mu;
                        }
                    }
                } else if (wave <= 4) {
                    if (tau > 0) apply(b ^ 1, sl, tau < n ? (ocur[tau] >> 2) : -1);
                    sl = sl + 1 == W ? 0 : sl + 1;
                } else {
                    // decision words of block kq + 1 (counter (replica's chain 0, sweep, site,
                    // TAG_WORLD)), evaluated 3+3+2+2 rounds over the first four steps of block kq
                    {
                        const int kq = tau / kb, q = tau - kq * kb;
                        if (q < 4 && (kq + 1) * kb < n) {
                            if (q == 0) {
                                const int rr = lane / kb, qs = lane - (lane / kb) * kb;
                                nxt.start(cbase + rr * P, sweep_id, static_cast<uint32_t>((kq + 1) * kb + qs), TAG_WORLD,
                                          a.k0, a.k1);
                                nxt.rounds<3>();
                            } else if (q == 1) {
                                nxt.rounds<3>();
                            } else if (q == 2) {
                                nxt.rounds<2>();
                            } else {
                                nxt.rounds<2>();
                                RW[((kq + 1) & 1) * 64 + lane] = nxt.c0;
                            }
                        }
                    }
                    // store the row fetched last step, shuffle, fetch this step's row
                    if (tau > 0) {
                        ring_store(tau - 1 + a.pfd, hslot);
                        hslot = hslot + 1 == W ? 0 : hslot + 1;
                    }
                    if (sh_ii > 0) {
                        // iteration sh_ii of shuffle_order(onext, ..., sweep_id + 1, ...)
                        const int it = n - 1 - sh_ii;
                        if ((it & 255) == 0) {
                            const int i0 = n - 1 - 4 * ((it >> 2) + lane);
                            if (i0 > 0)
                                shb = philox4x32_10(cbase, sweep_id + 1, static_cast<uint32_t>(i0), TAG_ORDER, a.k0, a.k1);
                        }
                        const int qq = it & 3;
                        const uint32_t wv = qq == 0 ? shb.x : qq == 1 ? shb.y : qq == 2 ? shb.z : shb.w;
                        const uint32_t d = __builtin_amdgcn_readlane(wv, (it >> 2) & 63);
                        if (lane == 0) {
                            const int j = static_cast<int>(below(d, static_cast<uint32_t>(sh_ii + 1)));
                            const int tmp = onext[sh_ii];
                            onext[sh_ii] = onext[j];
                            onext[j] = tmp;
                        }
                        --sh_ii;
                    }
                    ring_fetch(ocur, tau + a.pfd);
                }
                __syncthreads();
            }
            have_next = pre_next;
            if (pre_next) {
                next_id = sweep_id + 1;
                cur ^= 1;
            }
        }
        if (a.tail_mode != 2) group_tail(a, F, S, ldf, G, r0, g, step, L.ered, L.scal, gbest);
    }
    group_save(a, F, S, ldf, G, r0);
    if (tid == 0) a.accepted[g] = 0ull;  // worldline flips are not counted (as in the two-barrier kernel)
}

template <int NQ>
__global__ void __launch_bounds__(kThreads) sqa_pipe_wl_kernel(SqaArgs a) {
    pipe_wl_group<NQ>(a, static_cast<int>(blockIdx.x));
}

// ---- batched runs: persistent workgroups over (instance, replica group) items -----------------
struct BatchItem {
    int inst;   // index into the per-instance SqaArgs
    int group;  // replica group of that instance
};

__global__ void reset_counters_kernel(int *c, int count) {
    if (static_cast<int>(threadIdx.x) < count) c[threadIdx.x] = 0;
}

// V > 0: pipelined body with V 4-float row groups per lane; V == 0: two-barrier body with the
// state in LDS; V < 0: two-barrier body with the state in global memory. Every workgroup takes
// items from *counter until none is left (the exit every wave reaches).
template <int V>
__global__ void __launch_bounds__(kThreads) sqa_batch_kernel(const SqaArgs *__restrict__ inst,
                                                             const BatchItem *__restrict__ items, int nitems,
                                                             int *counter, int *done) {
    __shared__ int s_item;
    // at most nitems + 1 rounds: every round takes a fresh index from the counter
    for (int round = 0; round <= nitems; ++round) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const BatchItem b = items[it];
        if constexpr (V > 0) pipe_group<V>(inst[b.inst], b.group);
        else if constexpr (V == 0) general_group<true>(inst[b.inst], b.group);
        else general_group<false>(inst[b.inst], b.group);
        if (threadIdx.x == 0) atomicAdd(done, 1);  // items completed, checked by the host
    }
}

// ---- SQA-PT exchange (sqa_parallel_tempering.cpp:299-335), one workgroup per ladder ----------
// States stay where they are; the rung <-> state-slot map moves. For adjacent rungs a, b holding
// states with C = sum_t E_t and B = sum_{t,i} s_t,i s_t+1,i, the reference's 4-action difference
// (e_ab + e_ba) - (e_aa + e_bb) equals ((beta_a - beta_b)/P)(C_b - C_a) - (J_a - J_b)(B_b - B_a).
struct PtSwapArgs {
    int R, P, n, steps, step, do_swap;
    const double *chain_e;
    const int *chain_b;
    int *slot_of_rung;  // [L][R]
    int *rung_of_slot;  // [L*R]
    const double *bp;   // [R] beta / P
    const double *jpr;  // [R] J_perp
    uint32_t k0, k1;
    const int8_t *gS;
    double *best_e;     // [L]
    int8_t *best_state; // [L][n]
    double *avg_trace;  // [L][steps]
    double *swap_trace; // [L][steps]
};

__global__ void __launch_bounds__(64) pt_swap_kernel(PtSwapArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int l = blockIdx.x, tid = threadIdx.x, R = a.R, P = a.P;
    double *C = reinterpret_cast<double *>(smem);
    double *E = C + R;
    long long *B = reinterpret_cast<long long *>(E + R);
    int *src = reinterpret_cast<int *>(B + R);  // [0]: chain to copy into best_state, or -1
    int *sl = a.slot_of_rung + static_cast<int64_t>(l) * R;
    for (int r = tid; r < R; r += 64) {
        const int base = sl[r] * P;
        double cs = 0.0, mn = 1e308;
        long long bs = 0;
        for (int t = 0; t < P; ++t) {
            const double e = a.chain_e[base + t];
            cs += e;
            bs += a.chain_b[base + t];
            if (e < mn) mn = e;
        }
        C[r] = cs;
        E[r] = mn;
        B[r] = bs;
    }
    __syncthreads();
    if (tid == 0) {
        src[0] = -1;
        double be = a.best_e[l];
        for (int r = 0; r < R; ++r)
            if (E[r] < be) {
                be = E[r];
                const int base = sl[r] * P;
                int bt = 0;
                for (int t = 0; t < P; ++t)
                    if (a.chain_e[base + t] == E[r]) { bt = t; break; }
                src[0] = base + bt;
            }
        a.best_e[l] = be;
        double acc = 0.0, att = 0.0;
        if (a.do_swap) {
            for (int r = 0; r + 1 < R; ++r) {
                const double d = (a.bp[r] - a.bp[r + 1]) * (C[r + 1] - C[r]) -
                                 (a.jpr[r] - a.jpr[r + 1]) * static_cast<double>(B[r + 1] - B[r]);
                att += 1.0;
                bool ok = d <= 0.0;
                if (!ok) {
                    const Philox4 q = philox4x32_10(static_cast<uint32_t>(l), static_cast<uint32_t>(a.step),
                                                    static_cast<uint32_t>(r), 0x5A5A0000u, a.k0, a.k1);
                    ok = static_cast<double>(u01(q.x)) < exp(-d);
                }
                if (ok) {
                    int ts = sl[r]; sl[r] = sl[r + 1]; sl[r + 1] = ts;
                    double tc = C[r]; C[r] = C[r + 1]; C[r + 1] = tc;
                    double te = E[r]; E[r] = E[r + 1]; E[r + 1] = te;
                    long long tb = B[r]; B[r] = B[r + 1]; B[r + 1] = tb;
                    acc += 1.0;
                }
            }
        }
        double avg = 0.0;
        for (int r = 0; r < R; ++r) avg += E[r];
        a.avg_trace[static_cast<int64_t>(l) * a.steps + a.step] = avg / R;
        a.swap_trace[static_cast<int64_t>(l) * a.steps + a.step] = att > 0.0 ? acc / att : 0.0;
        for (int r = 0; r < R; ++r) a.rung_of_slot[sl[r]] = r;
    }
    __syncthreads();
    const int c = src[0];
    if (c >= 0)
        for (int i = tid; i < a.n; i += 64)
            a.best_state[static_cast<int64_t>(l) * a.n + i] = a.gS[static_cast<int64_t>(c) * a.n + i];
}

__global__ void gather_states(const int8_t *gS, const int *chain_of, int count, int n, int8_t *out) {
    const int k = blockIdx.x;
    if (k >= count) return;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        out[static_cast<int64_t>(k) * n + i] = gS[static_cast<int64_t>(chain_of[k]) * n + i];
}
}  // namespace

// Launch plan of the FP32 sweep kernel: replicas per workgroup, state home (LDS or global), J-row
// ring, dynamic LDS size.
constexpr bool n_fits_ring(int n) { return n <= 2048; }

struct SqaPlan {
    int64_t rg = 1, ngroups = 0;
    bool use_lds = false;
    int ring = 0, pfd = 1, ldf = 0, ldr = 0, pipe = 0;
    int hubo_lds = 0, hubo_nv = 0, adj_lds = 0;
    int cmax = kMaxChains;
    int s_lds = 0;  // global-state fields with the spins in LDS (dense float4 kernel)
    int la = 0;     // look-ahead slice sweep (s_lds, one wavefront of chains)
    int wl_ring = 0;  // LDS state without a slice row ring: rows of a worldline-only ring that fits
    int wlr = 0;      // this layout holds that worldline ring (a launch of worldline sweeps only)
    size_t shmem = 0;
};

// rg_force > 0 fixes the replicas per group (batch autotuner); lds_cap leaves room for a
// kernel's static LDS.
SqaPlan plan_sqa(const qa_ctx *ctx, const qa_problem *p, int64_t R, int64_t P, int64_t rg_force = 0,
                 size_t lds_cap = 160 * 1024) {
    const int64_t n = p->n;
    if (P > kMaxSlices)
        throw Error{QA_EUNSUPPORTED, "FP32 mode runs at most 384 Trotter slices per replica in this build"};
    if (p->kind == QA_KIND_DENSE && !p->symmetric)
        throw Error{QA_EUNSUPPORTED, "FP32 mode needs a symmetric J (the reference's field updates assume it); "
                                     "use the FP64 validation mode for an asymmetric J"};
    QA_REQUIRE(n < (1 << 24), "problem too large");
    SqaPlan pl;
    const int64_t rn = (n + 3) / 4 * 4;  // LDS row stride of fields and staged J rows
    // replicas per group: fill a 64-lane wave when there are enough replicas for every CU
    pl.rg = std::max<int64_t>(1, kMaxChains / P);
    while (pl.rg > 1 && (R + pl.rg - 1) / pl.rg < ctx->num_cus) --pl.rg;
    if (rg_force > 0) pl.rg = std::min<int64_t>(rg_force, std::max<int64_t>(1, kMaxChains / P));
    // per-chain scratch: one wavefront of entries, or P rounded up to whole wavefronts (rg = 1)
    pl.cmax = P <= kMaxChains ? kMaxChains : static_cast<int>((P + 63) / 64 * 64);
    // LDS without the row ring: two visit orders, per-chain scratch, and (LDS layout) the state
    auto fixed = [&](int64_t r, bool lds) {
        size_t b = 2 * lds_align16(static_cast<size_t>(n) * 4) + chain_scratch_bytes(pl.cmax);
        if (lds) b += lds_align16(static_cast<size_t>(r * P * rn) * 4) + lds_align16(static_cast<size_t>(r * P * n));
        return b;
    };
    // J-row ring slots beside an LDS state of r replicas: P + 2 .. P + 5 (prefetch distance 1..4)
    const bool dense_ring = p->kind == QA_KIND_DENSE && n_fits_ring(static_cast<int>(n));
    auto ring_rows = [&](int64_t r) -> int {
        if (!dense_ring) return 0;
        const size_t f = fixed(r, true);
        if (f + 16 >= lds_cap) return 0;
        const int64_t w = std::min<int64_t>(P + 5, static_cast<int64_t>((lds_cap - f - 16) / (static_cast<size_t>(rn) * 4)));
        return w >= P + 2 ? static_cast<int>(w) : 0;
    };
    while (pl.rg > 1 && (fixed(pl.rg, true) > lds_cap || (dense_ring && ring_rows(pl.rg) == 0))) --pl.rg;
    // QANNEAL_SQA_STATE=global keeps the state in global memory even when it fits LDS (the same
    // results bit for bit; a measurement and test switch)
    const char *state_env = std::getenv("QANNEAL_SQA_STATE");
    const bool force_global = state_env && std::strcmp(state_env, "global") == 0;
    pl.use_lds = fixed(pl.rg, true) <= lds_cap && !force_global;
    pl.ring = pl.use_lds ? ring_rows(pl.rg) : 0;
    pl.pfd = pl.ring ? pl.ring - static_cast<int>(P) - 1 : 1;
    const char *env = std::getenv("QANNEAL_SQA_PIPELINE");
    pl.pipe = (pl.use_lds && pl.ring > 0 && pl.cmax == kMaxChains && !(env && env[0] == '0')) ? 1 : 0;
    pl.ngroups = (R + pl.rg - 1) / pl.rg;
    pl.shmem = fixed(pl.rg, pl.use_lds) + lds_align16(static_cast<size_t>(pl.ring) * rn * 4);
    QA_REQUIRE(pl.shmem <= lds_cap, "problem too large for the FP32 sweep kernel's LDS layout");
    // worldline sweeps read one row per site (every slice at the same site), not one per slice:
    // where the slice ring does not fit beside the state, a ring of 3-8 rows may
    if (pl.use_lds && dense_ring && pl.ring == 0 && pl.cmax == kMaxChains) {
        const size_t f = fixed(pl.rg, true);
        const int64_t w = f + 16 < lds_cap ? std::min<int64_t>(8, static_cast<int64_t>((lds_cap - f - 16) / (static_cast<size_t>(rn) * 4))) : 0;
        if (w >= 3) pl.wl_ring = static_cast<int>(w);
    }
    // the coupling lists the field updates walk (sparse adjacency, HUBO incidence and terms) are
    // staged in LDS after everything above when they fit
    const char *lists_env = std::getenv("QANNEAL_SQA_LISTS_LDS");
    const bool lists = !(lists_env && lists_env[0] == '0');
    if (p->kind == QA_KIND_HUBO && lists) {
        const int nv = static_cast<int>(p->t_vars.count);
        const size_t hb = hubo_lds_bytes(static_cast<int>(n), static_cast<int>(p->nterms), nv);
        if (pl.shmem + hb <= lds_cap) {
            pl.hubo_lds = 1;
            pl.hubo_nv = nv;
            pl.shmem += hb;
        }
    }
    if (p->kind == QA_KIND_SPARSE && lists) {
        const size_t sb = adj_lds_bytes(static_cast<int>(n), static_cast<int>(2 * p->nedges));
        if (pl.shmem + sb <= lds_cap) {
            pl.adj_lds = 1;
            pl.shmem += sb;
        }
    }
    // global-state dense problems whose spins fit LDS beside the scratch keep them there (the
    // float4 kernel; QANNEAL_SQA_SPINS_LDS=0 keeps them in global memory: same results bit for bit)
    {
        const char *f4env = std::getenv("QANNEAL_SQA_GLOBAL_F4");
        const char *slenv = std::getenv("QANNEAL_SQA_SPINS_LDS");
        const size_t sb = lds_align16(static_cast<size_t>(pl.rg * P * n));
        if (!pl.use_lds && p->kind == QA_KIND_DENSE && n % 4 == 0 && n <= 8192 && !(f4env && f4env[0] == '0') &&
            !(slenv && slenv[0] == '0') && pl.shmem + sb <= lds_cap) {
            pl.s_lds = 1;
            pl.shmem += sb;
        }
        // with one wavefront of chains the slice sweeps take the look-ahead kernel
        // (QANNEAL_SQA_LOOKAHEAD=0: the one-step prefetch, the same results bit for bit)
        const char *laenv = std::getenv("QANNEAL_SQA_LOOKAHEAD");
        if (pl.s_lds && pl.cmax == kMaxChains && !(laenv && laenv[0] == '0') && pl.shmem + kLaBytes <= lds_cap) {
            pl.la = 1;
            pl.shmem += kLaBytes;
        }
    }
    pl.ldf = static_cast<int>(pl.use_lds ? rn : n);
    pl.ldr = static_cast<int>(rn);
    return pl;
}

// The layout of a launch of worldline sweeps alone for a plan whose slice ring does not fit: the
// same state with the plan's worldline ring (prefetch distance ring - 2).
SqaPlan wl_plan_of(const SqaPlan &pl) {
    SqaPlan w = pl;
    if (!pl.pipe && pl.wl_ring > 0) {
        w.ring = pl.wl_ring;
        w.pfd = pl.wl_ring - 2;
        w.wlr = 1;
        w.shmem = pl.shmem + lds_align16(static_cast<size_t>(pl.wl_ring) * pl.ldr * 4);
    }
    return w;
}

// The plan-derived fields of a launch (replica groups, LDS carve-up, row ring, staged lists): every
// launch, the batched one included, takes them from here.
void apply_plan(SqaArgs &a, const SqaPlan &pl) {
    a.hubo_lds = pl.hubo_lds;
    a.hubo_nv = pl.hubo_nv;
    a.adj_lds = pl.adj_lds;
    a.cmax = pl.cmax;
    a.s_lds = pl.s_lds;
    a.la = pl.la;
    a.ngroups = static_cast<int>(pl.ngroups);
    a.ldf = pl.ldf;
    a.ldr = pl.ldr;
    a.ring = pl.ring;
    a.pfd = pl.pfd;
}

void launch_sqa(const qa_ctx *ctx, const SqaPlan &pl, const SqaArgs &args) {
    SqaArgs a = args;
    apply_plan(a, pl);
    // the pipelined kernel runs plain slice sweeps over dense couplings with an LDS state
    const bool pipe = pl.pipe && a.cmax == kMaxChains && a.sweep_kind == 0 && a.wl == 0 && a.cl == 0 &&
                      a.pv.kind == QA_KIND_DENSE && a.bstat == nullptr;
    // worldline sweeps alone (the second launch of a split step, or wl-only runs) on the same
    // layout: the pipelined worldline kernel (QANNEAL_SQA_WL_PIPE=0: the two-barrier kernel)
    const char *wl_env = std::getenv("QANNEAL_SQA_WL_PIPE");
    const bool wl_pipe = (pl.pipe || pl.wlr) && a.cmax == kMaxChains && a.sweep_kind == 0 && a.spb == 0 &&
                         a.cl == 0 && a.wl > 0 && a.pv.kind == QA_KIND_DENSE && a.bstat == nullptr &&
                         64 / a.rg >= 4 && !(wl_env && wl_env[0] == '0');
    QA_REQUIRE(!pl.wlr || wl_pipe, "the worldline ring layout runs only the pipelined worldline kernel");
    const void *kern;
    unsigned threads = kThreads;
    if (pipe || wl_pipe) {
        QA_REQUIRE((pipe ? a.ring >= a.P + 2 : (a.pfd >= 1 && a.ring >= a.pfd + 2)) && a.ldf == a.ldr &&
                       a.ldf % 4 == 0 && n_fits_ring(a.pv.n),
                   "pipelined sweep needs the dense row ring");
        const int ngr = a.ldf / 4;
        if (pipe)
            kern = ngr <= 64 ? reinterpret_cast<const void *>(sqa_pipe_kernel<1>)
                 : ngr <= 128 ? reinterpret_cast<const void *>(sqa_pipe_kernel<2>)
                 : ngr <= 256 ? reinterpret_cast<const void *>(sqa_pipe_kernel<4>)
                              : reinterpret_cast<const void *>(sqa_pipe_kernel<8>);
        else
            kern = ngr <= 64 ? reinterpret_cast<const void *>(sqa_pipe_wl_kernel<1>)
                 : ngr <= 128 ? reinterpret_cast<const void *>(sqa_pipe_wl_kernel<2>)
                 : ngr <= 256 ? reinterpret_cast<const void *>(sqa_pipe_wl_kernel<4>)
                              : reinterpret_cast<const void *>(sqa_pipe_wl_kernel<8>);
    } else {
        const int n = a.pv.n;
        const char *env = std::getenv("QANNEAL_SQA_GLOBAL_F4");
        const bool f4ok = !pl.use_lds && a.pv.kind == QA_KIND_DENSE && n % 4 == 0 && n <= 8192 &&
                          !(env && env[0] == '0');
        QA_REQUIRE(!a.s_lds || f4ok, "spins in LDS need the float4 global-state kernel");
        if (pl.use_lds)
            kern = reinterpret_cast<const void *>(sqa_fp32_kernel<true>);
        else if (!f4ok)
            kern = reinterpret_cast<const void *>(sqa_fp32_kernel<false>);
        else if (a.s_lds && a.la && a.sweep_kind == 0 && a.wl == 0 && a.cl == 0 && a.bstat == nullptr)
            kern = n <= 1024 ? reinterpret_cast<const void *>(sqa_la_kernel<1>)
                 : n <= 2048 ? reinterpret_cast<const void *>(sqa_la_kernel<2>)
                 : n <= 4096 ? reinterpret_cast<const void *>(sqa_la_kernel<4>)
                             : reinterpret_cast<const void *>(sqa_la_kernel<8>);
        else if (a.s_lds)
            kern = n <= 1024 ? reinterpret_cast<const void *>(sqa_fp32_kernel<false, 1, true>)
                 : n <= 2048 ? reinterpret_cast<const void *>(sqa_fp32_kernel<false, 2, true>)
                 : n <= 4096 ? reinterpret_cast<const void *>(sqa_fp32_kernel<false, 4, true>)
                             : reinterpret_cast<const void *>(sqa_fp32_kernel<false, 8, true>);
        else if (n <= 1024)
            kern = reinterpret_cast<const void *>(sqa_fp32_kernel<false, 1>);
        else if (n <= 2048)
            kern = reinterpret_cast<const void *>(sqa_fp32_kernel<false, 2>);
        else if (n <= 4096)
            kern = reinterpret_cast<const void *>(sqa_fp32_kernel<false, 4>);
        else
            kern = reinterpret_cast<const void *>(sqa_fp32_kernel<false, 8>);
    }
    ensure_dynamic_lds(kern, pl.shmem);
    void *params[] = {&a};
    QA_HIP(hipLaunchKernel(kern, dim3(static_cast<unsigned>(pl.ngroups)), dim3(threads), params, pl.shmem,
                           ctx->stream));
    QA_HIP(hipGetLastError());
}

// A per-step observer: one launch per step on a device-resident state, the lattice copied back
// after each (SQAAnnealer::run's observer call, sqa_annealer.cpp:474-476).
void sqa_fp32_observed(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res) {
    if (prm->refresh_interval > 0)
        throw Error{QA_EUNSUPPORTED, "a per-step observer does not combine with refresh_interval"};
    const int64_t n = p->n, R = prm->replicas, P = prm->trotter_slices;
    SqaFp32Session f(ctx, p, R, P, prm->seed);
    std::vector<int8_t> h(R * P * n);
    for (int64_t step = 0; step < prm->steps; ++step) {
        const double bs = prm->betas[step] / static_cast<double>(P);
        const double jp = trotter_coupling(prm->betas[step], prm->gammas[step], P);
        f.run(1, &bs, &jp, 0, nullptr, prm->sweeps_per_beta, prm->cluster_sweeps, prm->worldline_sweeps, 0, nullptr,
              res->energy_trace + step);
        QA_HIP(hipMemcpyAsync(h.data(), f.device_spins(), R * P * n, hipMemcpyDeviceToHost, ctx->stream));
        QA_HIP(hipStreamSynchronize(ctx->stream));
        prm->on_step(prm->callback_user, step, res->energy_trace[step], h.data());
    }
    f.best(res->best_state, res->best_energy);
    if (res->kernel_ms) *res->kernel_ms = f.ms;
    if (res->accept_rate) *res->accept_rate = -1.0;
}

void sqa_fp32(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res) {
    if (prm->on_step) {
        sqa_fp32_observed(ctx, p, prm, res);
        return;
    }
    const int64_t n = p->n, R = prm->replicas, P = prm->trotter_slices;
    const SqaPlan pl = plan_sqa(ctx, p, R, P);
    const int64_t rg = pl.rg, ngroups = pl.ngroups;
    const bool use_lds = pl.use_lds;
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
    // launches of at most `chunk` schedule steps; between them the FP32 fields are refreshed
    const int64_t per_step = prm->sweeps_per_beta + prm->worldline_sweeps + prm->cluster_sweeps;
    const int64_t ri = prm->refresh_interval;
    QA_REQUIRE(ri >= 0, "refresh_interval must be >= 0");
    const int64_t chunk = ri > 0 ? std::max<int64_t>(1, ri / per_step) : prm->steps;
    const int64_t nchunks = (prm->steps + chunk - 1) / chunk;
    const int64_t nchain = R * P, nsample = std::min<int64_t>(nchain, 64);
    // Steps that mix slice sweeps with cluster or worldline sweeps run as two launches when the
    // slice sweeps qualify for the pipelined kernel: the slice sweeps there (no step tail), then the
    // cluster and worldline sweeps and the step tail in the two-barrier kernel, the LDS state
    // passed through gS/gF. Sweep ids, Philox counters and arithmetic are those of one launch.
    const char *split_env = std::getenv("QANNEAL_SQA_SPLIT");
    const char *wl_env = std::getenv("QANNEAL_SQA_WL_PIPE");
    // without the slice ring, steps split only for the pipelined worldline kernel on its own ring
    const bool wl_split = !pl.pipe && pl.wl_ring > 0 && p->kind == QA_KIND_DENSE && pl.rg <= 16 &&
                          prm->worldline_sweeps > 0 && prm->cluster_sweeps == 0 && !(wl_env && wl_env[0] == '0');
    // global-state fields with the spins in LDS: the slice sweeps of a step run in the look-ahead
    // kernel, the cluster and worldline sweeps and the tail in the one-step-prefetch kernel
    const bool la_split = !pl.use_lds && pl.s_lds && pl.la && p->kind == QA_KIND_DENSE;
    const bool split = (pl.pipe || wl_split || la_split) && prm->sweeps_per_beta > 0 &&
                       (prm->worldline_sweeps > 0 || prm->cluster_sweeps > 0) && !(split_env && split_env[0] == '0');
    const SqaPlan pl_rest = wl_split ? wl_plan_of(pl) : pl;
    const int64_t nacc_slots = split ? 2 * prm->steps : nchunks;
    DevBuf<int8_t> gS(R * P * n), bstate(ngroups * n);
    DevBuf<float> gF((use_lds && nchunks == 1 && !split) ? 1 : R * P * n);
    DevBuf<double> tpart(prm->steps * ngroups), be(ngroups);
    DevBuf<int> bstep(ngroups), bchain(ngroups);
    DevBuf<unsigned long long> accd(ngroups * nacc_slots), drift(2);
    const bool mfma_refresh = p->kind == QA_KIND_DENSE && p->symmetric;
    DevBuf<double> scratch(nchunks > 1 ? (mfma_refresh ? nsample : nchain) * n : 0);
    QA_HIP(hipMemsetAsync(drift.ptr, 0, 2 * sizeof(unsigned long long), s));
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

    QA_HIP(hipEventRecord(ctx->ev0, s));
    for (int64_t ci = 0; ci < nchunks; ++ci) {
        const int64_t start = ci * chunk, len = std::min(chunk, prm->steps - start);
        a.steps = static_cast<int>(len);
        a.bs = dbs.ptr + start;
        a.jp = djp.ptr + start;
        a.join = djoin.ptr + start;
        a.trace_part = tpart.ptr + start * ngroups;
        a.accepted = accd.ptr + ci * ngroups;
        a.resume = ci > 0 ? 1 : 0;
        a.save_fields = ci + 1 < nchunks ? 1 : 0;
        a.sweep_base = static_cast<uint32_t>(start * per_step);
        a.step_base = static_cast<int>(start);
        if (!split) {
            launch_sqa(ctx, pl, a);
        } else {
            for (int64_t k = start; k < start + len; ++k) {
                SqaArgs sl = a;  // slice sweeps of step k
                sl.steps = 1;
                sl.bs = dbs.ptr + k;
                sl.jp = djp.ptr + k;
                sl.join = djoin.ptr + k;
                sl.trace_part = tpart.ptr + k * ngroups;
                sl.wl = 0;
                sl.cl = 0;
                sl.tail_mode = 2;
                sl.resume = k > 0 ? 1 : 0;
                sl.save_fields = 1;
                sl.sweep_base = static_cast<uint32_t>(k * per_step);
                sl.step_base = static_cast<int>(k);
                sl.accepted = accd.ptr + 2 * k * ngroups;
                launch_sqa(ctx, pl, sl);
                SqaArgs rest = sl;  // cluster, then worldline sweeps, then the step tail
                rest.spb = 0;
                rest.wl = a.wl;
                rest.cl = a.cl;
                rest.tail_mode = 0;
                rest.resume = 1;
                rest.save_fields = k + 1 < prm->steps ? 1 : 0;
                rest.sweep_base = static_cast<uint32_t>(k * per_step + prm->sweeps_per_beta);
                rest.accepted = accd.ptr + (2 * k + 1) * ngroups;
                launch_sqa(ctx, pl_rest, rest);
            }
        }
        if (ci + 1 < nchunks) {  // drift of the running fields, matrix-core refresh, drift after it
            field_drift_launch(p, gS.ptr, gF.ptr, n, nsample, scratch.ptr, drift.ptr, s);
            fields_fp32_launch(p, gS.ptr, nchain, gF.ptr, n, scratch.ptr, s);
            field_drift_launch(p, gS.ptr, gF.ptr, n, nsample, scratch.ptr, drift.ptr + 1, s);
        }
    }
    QA_HIP(hipEventRecord(ctx->ev1, s));

    std::vector<double> hpart(prm->steps * ngroups), hbe(ngroups);
    std::vector<int> hstep(ngroups), hchain(ngroups);
    std::vector<unsigned long long> hacc(ngroups * nacc_slots), hdrift(2);
    tpart.download(hpart.data(), hpart.size(), s);
    be.download(hbe.data(), ngroups, s);
    bstep.download(hstep.data(), ngroups, s);
    bchain.download(hchain.data(), ngroups, s);
    accd.download(hacc.data(), ngroups * nacc_slots, s);
    drift.download(hdrift.data(), 2, s);
    QA_HIP(hipStreamSynchronize(s));
    if (res->field_drift) {
        const double nan = std::numeric_limits<double>::quiet_NaN();
        res->field_drift[0] = nchunks > 1 ? drift_value(hdrift[0]) : nan;
        res->field_drift[1] = nchunks > 1 ? drift_value(hdrift[1]) : nan;
    }
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


// ---- FP32 SQA run without any allocation (scratch in a static device array) ------------------
// The device memory the run needs beyond its problem and results -- schedule scalars, spins
// (and fields when they do not fit LDS), per-group best states, trace partials, counters, the
// reduced trace and the best sample -- is carved from one file-scope array. Each device holds its
// own copy; its address comes from hipGetSymbolAddress once per device and is used from then on
// (the array's name in host code is not a device address). Calls on one device are serialized by
// a host mutex; a call's contents are overwritten by the next call.
namespace {

constexpr size_t kNoAllocScratchBytes = size_t{64} << 20;  // 64 MiB per device
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxNoAllocDevices = 64;
std::mutex g_noalloc_mutex[kMaxNoAllocDevices];
unsigned char *g_noalloc_ptr[kMaxNoAllocDevices] = {};

// Set-up: the device address of this device's scratch (the device is current).
unsigned char *noalloc_scratch(int device) {
    QA_REQUIRE(device >= 0 && device < kMaxNoAllocDevices, "qa_sqa_run_noalloc: device index out of range");
    if (!g_noalloc_ptr[device]) {
        void *ptr = nullptr;
        QA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        QA_REQUIRE(ptr != nullptr, "qa_sqa_run_noalloc: no device address for the scratch array");
        g_noalloc_ptr[device] = static_cast<unsigned char *>(ptr);
    }
    return g_noalloc_ptr[device];
}

// Per-group best energies start at +inf and best chains at -1 (no hipMemsetAsync).
__global__ void noalloc_init_kernel(double *best_e, int *best_chain, int ngroups) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ngroups) {
        best_e[g] = __longlong_as_double(0x7FF0000000000000ll);
        best_chain[g] = -1;
    }
}

// Host-side reductions of sqa_fp32, on the device in the same order: the trace as the sequential
// sum over groups divided by R*P (one thread per step); the best group as the first strict
// minimum in (energy, step) over groups in order, its state copied to best_out; accepted moves
// summed. status[0] = best group or -1.
__global__ void noalloc_finish_kernel(const double *tpart, int steps, int ngroups, double rp, double *trace,
                                      const double *best_e, const int *best_step, const int *best_chain,
                                      const int8_t *bstate, int n, int8_t *best_out,
                                      const unsigned long long *acc, unsigned long long *acc_sum, int *status) {
    const int tid = threadIdx.x;
    for (int k = tid; k < steps; k += blockDim.x) {
        double sum = 0.0;
        for (int gi = 0; gi < ngroups; ++gi) sum += tpart[static_cast<int64_t>(k) * ngroups + gi];
        trace[k] = sum / rp;
    }
    __shared__ int s_bg;
    if (tid == 0) {
        int bg = -1;
        unsigned long long tot = 0;
        for (int gi = 0; gi < ngroups; ++gi) {
            tot += acc[gi];
            if (best_chain[gi] < 0) continue;
            if (bg < 0 || best_e[gi] < best_e[bg] || (best_e[gi] == best_e[bg] && best_step[gi] < best_step[bg]))
                bg = gi;
        }
        *acc_sum = tot;
        status[0] = bg;
        s_bg = bg;
    }
    __syncthreads();
    const int bg = s_bg;
    if (bg >= 0)
        for (int i = tid; i < n; i += blockDim.x) best_out[i] = bstate[static_cast<int64_t>(bg) * n + i];
}

// Bump allocator over the scratch array (256-byte aligned pieces).
struct Carve {
    unsigned char *base;
    size_t used = 0;
    template <class T>
    T *take(size_t count) {
        T *p = base ? reinterpret_cast<T *>(base + used) : nullptr;
        used += (count * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

}  // namespace

void sqa_fp32_noalloc(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res) {
    QA_REQUIRE(prm->mode == QA_MODE_FP32, "qa_sqa_run_noalloc: FP32 mode only");
    QA_REQUIRE(prm->refresh_interval == 0, "qa_sqa_run_noalloc: refresh_interval is not supported");
    QA_REQUIRE(!prm->on_step && !prm->on_sweep, "qa_sqa_run_noalloc: observer callbacks are not supported");
    const int64_t n = p->n, R = prm->replicas, P = prm->trotter_slices, steps = prm->steps;
    const SqaPlan pl = plan_sqa(ctx, p, R, P);
    const int64_t ngroups = pl.ngroups;
    // the same pieces sqa_fp32 allocates, sized first so that an oversized run is refused before
    // anything is launched
    auto need = [&](unsigned char *base) {
        Carve c{base};
        c.take<float>(steps);
        c.take<float>(steps);
        c.take<float>(steps);
        c.take<int8_t>(static_cast<size_t>(R * P * n));
        c.take<float>(pl.use_lds ? 1 : static_cast<size_t>(R * P * n));
        c.take<int8_t>(static_cast<size_t>(ngroups * n));
        c.take<double>(static_cast<size_t>(steps * ngroups));
        c.take<double>(ngroups);
        c.take<int>(ngroups);
        c.take<int>(ngroups);
        c.take<unsigned long long>(ngroups);
        c.take<double>(steps);
        c.take<int8_t>(n);
        c.take<double>(1);
        c.take<unsigned long long>(1);
        c.take<int>(1);
        return c.used;
    };
    const size_t bytes = need(nullptr);
    if (bytes > kNoAllocScratchBytes)
        throw Error{QA_EINVAL, "qa_sqa_run_noalloc: the run needs " + std::to_string(bytes) +
                                   " bytes of scratch, more than the " + std::to_string(kNoAllocScratchBytes) +
                                   " bytes this entry point holds per device (use qa_sqa_run)"};
    std::lock_guard<std::mutex> lock(g_noalloc_mutex[ctx->device]);
    Carve c{noalloc_scratch(ctx->device)};
    float *dbs = c.take<float>(steps), *djp = c.take<float>(steps), *djoin = c.take<float>(steps);
    int8_t *gS = c.take<int8_t>(static_cast<size_t>(R * P * n));
    float *gF = c.take<float>(pl.use_lds ? 1 : static_cast<size_t>(R * P * n));
    int8_t *bstate = c.take<int8_t>(static_cast<size_t>(ngroups * n));
    double *tpart = c.take<double>(static_cast<size_t>(steps * ngroups));
    double *be = c.take<double>(ngroups);
    int *bstep = c.take<int>(ngroups), *bchain = c.take<int>(ngroups);
    unsigned long long *accd = c.take<unsigned long long>(ngroups);
    double *dtrace = c.take<double>(steps);
    int8_t *best_out = c.take<int8_t>(n);
    double *dex = c.take<double>(1);
    unsigned long long *dacc = c.take<unsigned long long>(1);
    int *dstatus = c.take<int>(1);
    hipStream_t s = ctx->stream;

    // schedule scalars as sqa_fp32 computes them, staged through a small stack buffer
    constexpr int64_t kStage = 256;
    float stage[3][kStage];
    for (int64_t k0 = 0; k0 < steps; k0 += kStage) {
        const int64_t len = std::min(kStage, steps - k0);
        for (int64_t q = 0; q < len; ++q) {
            const int64_t k = k0 + q;
            const double jp = trotter_coupling(prm->betas[k], prm->gammas[k], P);
            stage[0][q] = static_cast<float>(prm->betas[k] / static_cast<double>(P));
            stage[1][q] = static_cast<float>(jp);
            stage[2][q] = static_cast<float>(1.0 - std::exp(-2.0 * jp));
        }
        QA_HIP(hipMemcpyAsync(dbs + k0, stage[0], len * sizeof(float), hipMemcpyHostToDevice, s));
        QA_HIP(hipMemcpyAsync(djp + k0, stage[1], len * sizeof(float), hipMemcpyHostToDevice, s));
        QA_HIP(hipMemcpyAsync(djoin + k0, stage[2], len * sizeof(float), hipMemcpyHostToDevice, s));
        QA_HIP(hipStreamSynchronize(s));  // the stack buffer is reused
    }
    hipLaunchKernelGGL(noalloc_init_kernel, dim3(static_cast<unsigned>((ngroups + 255) / 256)), dim3(256), 0, s, be,
                       bchain, static_cast<int>(ngroups));
    QA_HIP(hipGetLastError());

    SqaArgs a{};
    a.pv = view_of(p);
    a.P = static_cast<int>(P);
    a.R = static_cast<int>(R);
    a.rg = static_cast<int>(pl.rg);
    a.steps = static_cast<int>(steps);
    a.spb = static_cast<int>(prm->sweeps_per_beta);
    a.wl = static_cast<int>(prm->worldline_sweeps);
    a.cl = static_cast<int>(prm->cluster_sweeps);
    a.bs = dbs;
    a.jp = djp;
    a.join = djoin;
    a.k0 = static_cast<uint32_t>(prm->seed);
    a.k1 = static_cast<uint32_t>(prm->seed >> 32);
    a.gS = gS;
    a.gF = gF;
    a.trace_part = tpart;
    a.best_e = be;
    a.best_step = bstep;
    a.best_chain = bchain;
    a.best_state = bstate;
    a.accepted = accd;
    QA_HIP(hipEventRecord(ctx->ev0, s));
    launch_sqa(ctx, pl, a);
    QA_HIP(hipEventRecord(ctx->ev1, s));
    hipLaunchKernelGGL(noalloc_finish_kernel, dim3(1), dim3(256), 0, s, tpart, static_cast<int>(steps),
                       static_cast<int>(ngroups), static_cast<double>(R * P), dtrace, be, bstep, bchain, bstate,
                       static_cast<int>(n), best_out, accd, dacc, dstatus);
    QA_HIP(hipGetLastError());
    device_energies(p, best_out, 1, dex, s);
    int status = -1;
    unsigned long long tot = 0;
    QA_HIP(hipMemcpyAsync(&status, dstatus, sizeof(int), hipMemcpyDeviceToHost, s));
    QA_HIP(hipMemcpyAsync(&tot, dacc, sizeof(tot), hipMemcpyDeviceToHost, s));
    QA_HIP(hipMemcpyAsync(res->energy_trace, dtrace, steps * sizeof(double), hipMemcpyDeviceToHost, s));
    QA_HIP(hipMemcpyAsync(res->best_state, best_out, n, hipMemcpyDeviceToHost, s));
    QA_HIP(hipMemcpyAsync(res->best_energy, dex, sizeof(double), hipMemcpyDeviceToHost, s));
    QA_HIP(hipStreamSynchronize(s));
    QA_REQUIRE(status >= 0, "no finite slice energy was produced");
    float ms = 0;
    QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (res->kernel_ms) *res->kernel_ms = ms;
    if (res->accept_rate) {
        const double att = static_cast<double>(steps) * prm->sweeps_per_beta * R * P * n;
        *res->accept_rate = att > 0 ? static_cast<double>(tot) / att : 0.0;
    }
    if (res->field_drift) res->field_drift[0] = res->field_drift[1] = std::numeric_limits<double>::quiet_NaN();
}


// ---- device-resident session for the adaptive J_perp schedules ------------------------------
SqaFp32Session::SqaFp32Session(qa_ctx *ctx, const qa_problem *p, int64_t R, int64_t P, uint64_t key)
    : ctx_(ctx), p_(p), R_(R), P_(P), key_(key) {
    const SqaPlan pl = plan_sqa(ctx, p, R, P);
    rg_ = pl.rg;
    ngroups_ = pl.ngroups;
    use_lds_ = pl.use_lds;
    const int64_t n = p->n;
    gS_.alloc(R * P * n);
    gF_.alloc(R * P * n);
    bstate_.alloc(ngroups_ * n);
    be_.alloc(ngroups_);
    bstep_.alloc(ngroups_);
    bchain_.alloc(ngroups_);
    std::vector<double> inf(ngroups_, std::numeric_limits<double>::infinity());
    std::vector<int> neg(ngroups_, -1);
    be_.upload(inf.data(), ngroups_, ctx->stream);
    bchain_.upload(neg.data(), ngroups_, ctx->stream);
    QA_HIP(hipStreamSynchronize(ctx->stream));
}

void SqaFp32Session::copy_state_from(const SqaFp32Session &o) {
    QA_REQUIRE(o.R_ == R_ && o.P_ == P_ && o.p_ == p_ && o.started_, "session shapes differ");
    const size_t cnt = static_cast<size_t>(R_ * P_ * p_->n);
    QA_HIP(hipMemcpyAsync(gS_.ptr, o.gS_.ptr, cnt, hipMemcpyDeviceToDevice, ctx_->stream));
    QA_HIP(hipMemcpyAsync(gF_.ptr, o.gF_.ptr, cnt * sizeof(float), hipMemcpyDeviceToDevice, ctx_->stream));
    started_ = true;
}

void SqaFp32Session::save() {
    QA_REQUIRE(started_, "nothing to save");
    const size_t cnt = static_cast<size_t>(R_ * P_ * p_->n);
    if (!sS_.ptr) {
        sS_.alloc(cnt);
        sF_.alloc(cnt);
    }
    QA_HIP(hipMemcpyAsync(sS_.ptr, gS_.ptr, cnt, hipMemcpyDeviceToDevice, ctx_->stream));
    QA_HIP(hipMemcpyAsync(sF_.ptr, gF_.ptr, cnt * sizeof(float), hipMemcpyDeviceToDevice, ctx_->stream));
}

void SqaFp32Session::restore() {
    QA_REQUIRE(sS_.ptr, "no saved state");
    const size_t cnt = static_cast<size_t>(R_ * P_ * p_->n);
    QA_HIP(hipMemcpyAsync(gS_.ptr, sS_.ptr, cnt, hipMemcpyDeviceToDevice, ctx_->stream));
    QA_HIP(hipMemcpyAsync(gF_.ptr, sF_.ptr, cnt * sizeof(float), hipMemcpyDeviceToDevice, ctx_->stream));
}

void SqaFp32Session::run(int64_t k, const double *bs, const double *jp, int64_t nparam, const int *param_of_rep,
                         int64_t spb, int64_t cl, int64_t wl, int tail, BondStats *stats, double *trace) {
    QA_REQUIRE(k > 0, "run: no steps");
    QA_REQUIRE(!param_of_rep || k == 1, "per-replica parameters run one step per launch");
    const SqaPlan pl = plan_sqa(ctx_, p_, R_, P_);
    hipStream_t s = ctx_->stream;
    const int64_t np = param_of_rep ? nparam : k;
    std::vector<float> hbs(np), hjp(np), hjoin(np);
    for (int64_t q = 0; q < np; ++q) {
        hbs[q] = static_cast<float>(bs[q]);
        hjp[q] = static_cast<float>(jp[q]);
        hjoin[q] = static_cast<float>(1.0 - std::exp(-2.0 * jp[q]));
    }
    DevBuf<float> dbs(np), djp(np), djoin(np);
    DevBuf<int> dpor(param_of_rep ? R_ : 0);
    dbs.upload(hbs.data(), np, s);
    djp.upload(hjp.data(), np, s);
    djoin.upload(hjoin.data(), np, s);
    dpor.upload(param_of_rep, param_of_rep ? R_ : 0, s);
    DevBuf<double> bst(stats ? k * ngroups_ * 2 : 0), bend(stats ? k : 0), tpart(tail == 0 ? k * ngroups_ : 0);
    DevBuf<unsigned long long> acc(ngroups_);
    if (tail == 1 && !ce_.ptr) {
        ce_.alloc(R_ * P_);
        cb_.alloc(R_ * P_);
    }
    SqaArgs a{};
    a.pv = view_of(p_);
    a.P = static_cast<int>(P_);
    a.R = static_cast<int>(R_);
    a.rg = static_cast<int>(rg_);
    a.steps = static_cast<int>(k);
    a.spb = static_cast<int>(spb);
    a.wl = static_cast<int>(wl);
    a.cl = static_cast<int>(cl);
    a.bs = dbs.ptr;
    a.jp = djp.ptr;
    a.join = djoin.ptr;
    a.k0 = static_cast<uint32_t>(key_);
    a.k1 = static_cast<uint32_t>(key_ >> 32);
    a.gS = gS_.ptr;
    a.gF = gF_.ptr;
    a.trace_part = tpart.ptr;
    a.best_e = be_.ptr;
    a.best_step = bstep_.ptr;
    a.best_chain = bchain_.ptr;
    a.best_state = bstate_.ptr;
    a.accepted = acc.ptr;
    a.param_of_rep = dpor.ptr;
    a.resume = started_ ? 1 : 0;
    a.save_fields = 1;
    a.tail_mode = tail;
    a.chain_e = ce_.ptr;
    a.chain_b = cb_.ptr;
    a.sweep_base = sweeps_;
    a.step_base = static_cast<int>(recorded_);
    a.bstat = bst.ptr;
    a.bend = bend.ptr;
    QA_HIP(hipEventRecord(ctx_->ev0, s));
    launch_sqa(ctx_, pl, a);
    QA_HIP(hipEventRecord(ctx_->ev1, s));
    started_ = true;
    sweeps_ += static_cast<uint32_t>(k * (spb + cl + wl));
    if (tail == 0) recorded_ += k;
    std::vector<double> hb(bst.count), he(bend.count), ht(tpart.count);
    bst.download(hb.data(), hb.size(), s);
    bend.download(he.data(), he.size(), s);
    tpart.download(ht.data(), ht.size(), s);
    if (tail == 1) {
        chain_e.resize(R_ * P_);
        chain_b.resize(R_ * P_);
        ce_.download(chain_e.data(), R_ * P_, s);
        cb_.download(chain_b.data(), R_ * P_, s);
    }
    QA_HIP(hipStreamSynchronize(s));
    float t = 0.0f;
    QA_HIP(hipEventElapsedTime(&t, ctx_->ev0, ctx_->ev1));
    ms += t;
    for (int64_t q = 0; q < k; ++q) {
        if (stats) {
            BondStats &st = stats[q];
            st = BondStats{};
            for (int64_t g = 0; g < ngroups_; ++g) {
                st.sum += hb[(q * ngroups_ + g) * 2];
                st.sum2 += hb[(q * ngroups_ + g) * 2 + 1];
            }
            st.count = R_ * (spb + cl);
            st.b0 = he[q];
        }
        if (tail == 0 && trace) {
            double sum = 0.0;
            for (int64_t g = 0; g < ngroups_; ++g) sum += ht[q * ngroups_ + g];
            trace[q] = sum / static_cast<double>(R_ * P_);
        }
    }
}

void SqaFp32Session::best(int8_t *state, double *energy) {
    hipStream_t s = ctx_->stream;
    std::vector<double> hbe(ngroups_);
    std::vector<int> hstep(ngroups_), hchain(ngroups_);
    be_.download(hbe.data(), ngroups_, s);
    bstep_.download(hstep.data(), ngroups_, s);
    bchain_.download(hchain.data(), ngroups_, s);
    QA_HIP(hipStreamSynchronize(s));
    int64_t bg = -1;  // first strict minimum in (step, replica, slice) order
    for (int64_t gi = 0; gi < ngroups_; ++gi) {
        if (hchain[gi] < 0) continue;
        if (bg < 0 || hbe[gi] < hbe[bg] || (hbe[gi] == hbe[bg] && hstep[gi] < hstep[bg])) bg = gi;
    }
    QA_REQUIRE(bg >= 0, "no finite slice energy was produced");
    const int64_t n = p_->n;
    QA_HIP(hipMemcpyAsync(state, bstate_.ptr + bg * n, n, hipMemcpyDeviceToHost, s));
    DevBuf<double> ex(1);
    device_energies(p_, bstate_.ptr + bg * n, 1, ex.ptr, s);
    ex.download(energy, 1, s);
    QA_HIP(hipStreamSynchronize(s));
}

void sqapt_fp32(qa_ctx *ctx, const qa_problem *p, const qa_sqapt_params *prm, qa_sqapt_result *res) {
    QA_REQUIRE(res->final_states && res->final_energies && res->best_state && res->best_energy &&
                   res->average_energy_trace && res->swap_acceptance_trace, "qa_sqapt_run: NULL result array");
    const int64_t n = p->n, R = prm->rungs, L = prm->ladders, P = prm->trotter_slices, steps = prm->steps;
    const int64_t NS = L * R;  // state slots, one replica each
    const SqaPlan pl = plan_sqa(ctx, p, NS, P);
    hipStream_t s = ctx->stream;
    // ladder rows: one (fixed ladder) or one per step (step_betas / step_gammas), R pairs each
    const int64_t rows = prm->step_betas ? steps : 1, RL = rows * R;
    const double *lb = prm->step_betas ? prm->step_betas : prm->betas;
    const double *lg = prm->step_gammas ? prm->step_gammas : prm->gammas;
    std::vector<float> hbs(RL), hjp(RL), hjoin(RL);
    std::vector<double> hbp(RL), hjpd(RL);
    for (int64_t r = 0; r < RL; ++r) {
        hjpd[r] = trotter_coupling(lb[r], lg[r], P);
        hbp[r] = lb[r] / static_cast<double>(P);
        hbs[r] = static_cast<float>(hbp[r]);
        hjp[r] = static_cast<float>(hjpd[r]);
        hjoin[r] = static_cast<float>(1.0 - std::exp(-2.0 * hjpd[r]));
    }
    std::vector<int> sor(NS), ros(NS);
    for (int64_t l = 0; l < L; ++l)
        for (int64_t r = 0; r < R; ++r) {
            sor[l * R + r] = static_cast<int>(l * R + r);
            ros[l * R + r] = static_cast<int>(r);
        }
    DevBuf<float> dbs(RL), djp(RL), djoin(RL), gF(NS * P * n);
    DevBuf<double> dbp(RL), djpd(RL), chain_e(NS * P), be(L), at(L * steps), st(L * steps);
    DevBuf<int> chain_b(NS * P), dsor(NS), dros(NS);
    DevBuf<int8_t> gS(NS * P * n), bstate(L * n);
    // accepted moves per replica group: one slot per launch when the caller asks for the rate
    DevBuf<unsigned long long> accd(pl.ngroups * (res->accept_rate ? steps : 1)), drift(2);
    const int64_t ri = prm->refresh_interval;
    QA_REQUIRE(ri >= 0, "refresh_interval must be >= 0");
    const int64_t nchain = NS * P, nsample = std::min<int64_t>(nchain, 64);
    const bool mfma_refresh = p->kind == QA_KIND_DENSE && p->symmetric;
    DevBuf<double> scratch(ri > 0 ? (mfma_refresh ? nsample : nchain) * n : 0);
    QA_HIP(hipMemsetAsync(drift.ptr, 0, 2 * sizeof(unsigned long long), s));
    dbs.upload(hbs.data(), RL, s);
    djp.upload(hjp.data(), RL, s);
    djoin.upload(hjoin.data(), RL, s);
    dbp.upload(hbp.data(), RL, s);
    djpd.upload(hjpd.data(), RL, s);
    dsor.upload(sor.data(), NS, s);
    dros.upload(ros.data(), NS, s);
    std::vector<double> inf(L, std::numeric_limits<double>::infinity());
    be.upload(inf.data(), L, s);

    SqaArgs a{};
    a.pv = view_of(p);
    a.P = static_cast<int>(P);
    a.R = static_cast<int>(NS);
    a.rg = static_cast<int>(pl.rg);
    a.steps = 1;
    a.spb = static_cast<int>(prm->sweeps_per_step);
    a.wl = static_cast<int>(prm->worldline_sweeps);
    a.cl = static_cast<int>(prm->cluster_sweeps);
    a.bs = dbs.ptr;
    a.jp = djp.ptr;
    a.join = djoin.ptr;
    a.k0 = static_cast<uint32_t>(prm->seed);
    a.k1 = static_cast<uint32_t>(prm->seed >> 32);
    a.gS = gS.ptr;
    a.gF = gF.ptr;
    a.accepted = accd.ptr;
    a.param_of_rep = dros.ptr;
    a.save_fields = 1;
    a.tail_mode = 1;
    a.chain_e = chain_e.ptr;
    a.chain_b = chain_b.ptr;
    PtSwapArgs w{};
    w.R = static_cast<int>(R);
    w.P = static_cast<int>(P);
    w.n = static_cast<int>(n);
    w.steps = static_cast<int>(steps);
    w.chain_e = chain_e.ptr;
    w.chain_b = chain_b.ptr;
    w.slot_of_rung = dsor.ptr;
    w.rung_of_slot = dros.ptr;
    w.bp = dbp.ptr;
    w.jpr = djpd.ptr;
    w.k0 = a.k0 ^ 0x13579BDFu;
    w.k1 = a.k1;
    w.gS = gS.ptr;
    w.best_e = be.ptr;
    w.best_state = bstate.ptr;
    w.avg_trace = at.ptr;
    w.swap_trace = st.ptr;
    const size_t wsh = static_cast<size_t>(R) * 24 + 16;
    const uint32_t per_step = static_cast<uint32_t>(prm->sweeps_per_step + prm->worldline_sweeps + prm->cluster_sweeps);
    int64_t refreshes = 0;
    // optional: device time of the sweep launches alone (an event pair around each)
    std::vector<hipEvent_t> sev;
    if (res->sweep_ms) {
        sev.resize(static_cast<size_t>(2 * steps));
        for (auto &e : sev) QA_HIP(hipEventCreate(&e));
    }
    QA_HIP(hipEventRecord(ctx->ev0, s));
    for (int64_t step = 0; step < steps; ++step) {
        a.resume = step > 0 ? 1 : 0;
        a.sweep_base = static_cast<uint32_t>(step) * per_step;
        a.step_base = static_cast<int>(step);
        if (rows > 1) {
            a.bs = dbs.ptr + step * R;
            a.jp = djp.ptr + step * R;
            a.join = djoin.ptr + step * R;
            w.bp = dbp.ptr + step * R;
            w.jpr = djpd.ptr + step * R;
        }
        if (res->accept_rate) a.accepted = accd.ptr + step * pl.ngroups;
        if (res->sweep_ms) QA_HIP(hipEventRecord(sev[2 * step], s));
        launch_sqa(ctx, pl, a);
        if (res->sweep_ms) QA_HIP(hipEventRecord(sev[2 * step + 1], s));
        w.step = static_cast<int>(step);
        w.do_swap = ((step + 1) % prm->swap_interval == 0) ? 1 : 0;
        hipLaunchKernelGGL(pt_swap_kernel, dim3(static_cast<unsigned>(L)), dim3(64), wsh, s, w);
        QA_HIP(hipGetLastError());
        // every refresh_interval sweeps: drift of the running fields, matrix-core refresh, drift after
        if (ri > 0 && step + 1 < steps && ((step + 1) * static_cast<int64_t>(per_step)) / ri > refreshes) {
            refreshes = ((step + 1) * static_cast<int64_t>(per_step)) / ri;
            field_drift_launch(p, gS.ptr, gF.ptr, n, nsample, scratch.ptr, drift.ptr, s);
            fields_fp32_launch(p, gS.ptr, nchain, gF.ptr, n, scratch.ptr, s);
            field_drift_launch(p, gS.ptr, gF.ptr, n, nsample, scratch.ptr, drift.ptr + 1, s);
        }
    }
    QA_HIP(hipEventRecord(ctx->ev1, s));
    if (res->field_drift) {
        std::vector<unsigned long long> hdrift(2);
        drift.download(hdrift.data(), 2, s);
        QA_HIP(hipStreamSynchronize(s));
        const double nan = std::numeric_limits<double>::quiet_NaN();
        res->field_drift[0] = refreshes > 0 ? drift_value(hdrift[0]) : nan;
        res->field_drift[1] = refreshes > 0 ? drift_value(hdrift[1]) : nan;
    }
    std::vector<double> he(NS * P);
    chain_e.download(he.data(), NS * P, s);
    dsor.download(sor.data(), NS, s);
    QA_HIP(hipStreamSynchronize(s));
    // final states: per rung the projected (lowest-energy) slice of the state it holds
    std::vector<int> pick(NS);
    for (int64_t k = 0; k < NS; ++k) {
        const int base = sor[k] * static_cast<int>(P);
        int bt = 0;
        for (int t = 1; t < P; ++t)
            if (he[base + t] < he[base + bt]) bt = t;
        pick[k] = base + bt;
    }
    DevBuf<int> dpick(NS);
    DevBuf<int8_t> fs(NS * n);
    DevBuf<double> fe(NS), bex(L);
    dpick.upload(pick.data(), NS, s);
    hipLaunchKernelGGL(gather_states, dim3(static_cast<unsigned>(NS)), dim3(256), 0, s, gS.ptr, dpick.ptr,
                       static_cast<int>(NS), static_cast<int>(n), fs.ptr);
    QA_HIP(hipGetLastError());
    device_energies(p, fs.ptr, NS, fe.ptr, s);
    device_energies(p, bstate.ptr, L, bex.ptr, s);
    fs.download(res->final_states, NS * n, s);
    fe.download(res->final_energies, NS, s);
    bstate.download(res->best_state, L * n, s);
    bex.download(res->best_energy, L, s);
    at.download(res->average_energy_trace, L * steps, s);
    st.download(res->swap_acceptance_trace, L * steps, s);
    QA_HIP(hipStreamSynchronize(s));
    // a ladder whose best was a final-state slice keeps the lower of the two exact energies
    for (int64_t l = 0; l < L; ++l)
        for (int64_t r = 0; r < R; ++r)
            if (res->final_energies[l * R + r] < res->best_energy[l]) {
                res->best_energy[l] = res->final_energies[l * R + r];
                std::memcpy(res->best_state + l * n, res->final_states + (l * R + r) * n, n);
            }
    if (res->swap_uniforms_used) *res->swap_uniforms_used = 0;
    if (res->kernel_ms) {
        float ms = 0;
        QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *res->kernel_ms = ms;
    }
    if (res->sweep_ms) {
        double tot = 0.0;
        for (int64_t step = 0; step < steps; ++step) {
            float ms = 0;
            QA_HIP(hipEventElapsedTime(&ms, sev[2 * step], sev[2 * step + 1]));
            tot += ms;
        }
        for (auto &e : sev) (void)hipEventDestroy(e);
        *res->sweep_ms = tot;
    }
    if (res->accept_rate) {
        std::vector<unsigned long long> hacc(pl.ngroups * steps);
        accd.download(hacc.data(), hacc.size(), s);
        QA_HIP(hipStreamSynchronize(s));
        double tot = 0.0;
        for (unsigned long long v : hacc) tot += static_cast<double>(v);
        *res->accept_rate = tot / (static_cast<double>(steps) * per_step * NS * P * n);
    }
}


namespace {
// ---- sqa_chi checkerboard sweeps (pilot and production): one wavefront per chain --------------
// In a checkerboard phase (sqa_chi_annealer.cpp:81-140) every chain of the phase's colour sweeps
// sites 0..n-1 in order while the chains of the other colours are held fixed, so the chains of a
// phase are independent and each one's sites depend only on its own flips. One launch per phase,
// one wavefront per chain: the chain's fields and spins, the fixed neighbour sums and the phase's
// Philox words sit in the wave's LDS; a dense J row is loaded for an accepted flip (kDemand, the
// default) or streamed in site order with the loads issued two sites ahead (register sets A and
// B, the site loop unrolled by two); the flip updates the fields with the same fmaf per element,
// in the same site order, as the two-barrier kernel's checkerboard sweep: the trajectory is the
// same bit for bit. Sparse problems stage their adjacency (CSR) in the wave's LDS instead; a
// flip's neighbour updates run one per lane when no adjacency row lists a spin twice, else one
// after another on lane 0, in row order.
struct ChkArgs {
    const float *J4rows;  // dense: J32p, n rows of rn floats (rn = n rounded up to 4)
    const int32_t *adj_ptr, *adj_idx;  // sparse: CSR [n + 1], [nnz]
    const float *adj_val;              // sparse: [nnz]
    const int32_t *inc_ptr, *inc_idx, *t_off, *t_vars;  // HUBO: incidence CSR and term lists
    const float *t_coef;                                // HUBO: [nterms]
    int8_t *gS;           // [R*P][n]
    float *gF;            // [R*P][n]
    const float *bsv, *jpv;  // beta / P and J_perp by parameter index
    const int *pof;          // non-null: replica r uses index pof[r] (pilot scan points), else pidx
    int pidx;
    int P, n, rn, ph, nnz, unique, nterms, nv;
    int lists_lds;           // sparse / HUBO: 1 = coupling lists staged in the wave's LDS
    uint32_t sweep_id, k0, k1;
};

// Random words of a chunk of sites: 64 Philox blocks, one per lane (>= 253 consecutive sites).
constexpr int kChkWords = 256;

// LDS bytes of chk_wave_kernel: fields, a chunk of random words, spins, neighbour sums
__host__ __device__ constexpr size_t chk_lds_bytes(int n, int rn) {
    return lds_align16(static_cast<size_t>(rn) * 4 + kChkWords * 4 + static_cast<size_t>(n) * 2);
}
// ... plus the coupling lists when staged: sparse CSR, or the HUBO arrays
__host__ __device__ constexpr size_t chk_list_bytes(int kind, int n, int nnz, int nterms, int nv) {
    return kind == QA_KIND_SPARSE ? lds_align16(static_cast<size_t>(n + 1) * 4) + static_cast<size_t>(nnz) * 8
         : kind == QA_KIND_HUBO   ? hubo_lds_bytes(n, nterms, nv)
                                  : 0;
}

// kDemand (dense): load a J row only for an accepted flip instead of streaming every row ahead of
// its site; with thousands of chains in flight (the pilot scan) this cuts the row traffic to the
// accepted fraction while other waves hide each load's latency.
// NQ > 0: dense, float4 groups of a row per lane (rn <= 256 * NQ); 0: sparse; -1: HUBO.
template <int NQ, bool kDemand = false>
__global__ void __launch_bounds__(64) chk_wave_kernel(ChkArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = blockIdx.x, lane = threadIdx.x;
    const int P = a.P, n = a.n, rn = a.rn, ng = a.rn >> 2;
    const int r = c / P, t = c - r * P;
    const int colour = P == 1 ? 0 : ((P & 1) ? (t == P - 1 ? 1 : ((t & 1) ? 2 : 0)) : (t & 1));
    if (colour != a.ph) return;  // the whole wave: this kernel has no workgroup barrier after here
    const int cprev = r * P + (t == 0 ? P - 1 : t - 1), cnext = r * P + (t + 1 == P ? 0 : t + 1);
    const int pq = a.pof ? a.pof[r] : a.pidx;
    const float bs = a.bsv[pq], jp = a.jpv[pq];
    float *F = reinterpret_cast<float *>(smem);           // [rn]
    f4 *F4 = reinterpret_cast<f4 *>(smem);
    uint32_t *W = reinterpret_cast<uint32_t *>(F + rn);   // [kChkWords] random words of the current chunk
    int8_t *S = reinterpret_cast<int8_t *>(W + kChkWords);  // [n]
    int8_t *NB = S + n;                                   // [n] s_{t-1,i} + s_{t+1,i} (fixed in the phase)
    unsigned char *lists = smem + chk_lds_bytes(n, rn);
    // sparse CSR and HUBO arrays: LDS copies (staged below) or the global arrays
    const int32_t *AP = a.adj_ptr, *AI = a.adj_idx, *IP = a.inc_ptr, *II = a.inc_idx, *TO = a.t_off, *TV = a.t_vars;
    const float *AV = a.adj_val, *TC = a.t_coef;
    if constexpr (NQ == 0) {
        if (a.lists_lds) {
            int32_t *ap = reinterpret_cast<int32_t *>(lists);
            int32_t *ai = ap + lds_align16(static_cast<size_t>(n + 1) * 4) / 4;
            float *av = reinterpret_cast<float *>(ai + a.nnz);
            for (int i = lane; i <= n; i += 64) ap[i] = a.adj_ptr[i];
            for (int q = lane; q < a.nnz; q += 64) {
                ai[q] = a.adj_idx[q];
                av[q] = a.adj_val[q];
            }
            AP = ap;
            AI = ai;
            AV = av;
        }
    } else if constexpr (NQ < 0) {
        if (a.lists_lds) {  // the layout of carve_lds's staged HUBO arrays
            unsigned char *q = lists;
            int32_t *ip = reinterpret_cast<int32_t *>(q);
            q += lds_align16(static_cast<size_t>(n + 1) * 4);
            int32_t *ii = reinterpret_cast<int32_t *>(q);
            q += lds_align16(static_cast<size_t>(a.nv) * 4);
            int32_t *to = reinterpret_cast<int32_t *>(q);
            q += lds_align16(static_cast<size_t>(a.nterms + 1) * 4);
            int32_t *tv = reinterpret_cast<int32_t *>(q);
            q += lds_align16(static_cast<size_t>(a.nv) * 4);
            float *tc = reinterpret_cast<float *>(q);
            for (int e = lane; e <= n; e += 64) ip[e] = a.inc_ptr[e];
            for (int e = lane; e < a.nv; e += 64) {
                ii[e] = a.inc_idx[e];
                tv[e] = a.t_vars[e];
            }
            for (int e = lane; e <= a.nterms; e += 64) to[e] = a.t_off[e];
            for (int e = lane; e < a.nterms; e += 64) tc[e] = a.t_coef[e];
            IP = ip;
            II = ii;
            TO = to;
            TV = tv;
            TC = tc;
        }
    }
    const int64_t off = static_cast<int64_t>(c) * n;
    for (int i = lane; i < n; i += 64) {
        S[i] = a.gS[off + i];
        NB[i] = static_cast<int8_t>(a.gS[static_cast<int64_t>(cprev) * n + i] + a.gS[static_cast<int64_t>(cnext) * n + i]);
        F[i] = a.gF[off + i];
    }
    for (int i = n + lane; i < rn; i += 64) F[i] = 0.0f;
    constexpr int NR = NQ > 0 ? NQ : 1;
    const f4 *J4 = reinterpret_cast<const f4 *>(a.J4rows);
    auto load_row = [&](f4 (&h)[NR], int k) {
        const f4 *row = J4 + static_cast<int64_t>(min(k, n - 1)) * ng;
        static_for<(NQ > 0 ? NQ : 0)>([&](auto q) { h[q] = row[min(lane + 64 * q, ng - 1)]; });
    };
    auto site = [&](const f4 (&h)[NR], int k, uint32_t word) {
        const float f = F[k];
        const int8_t s = S[k];
        const int nb = NB[k];
        const float fs = static_cast<float>(s);
        const float de = fmaf(bs, -2.0f * fs * f, 2.0f * jp * fs * static_cast<float>(nb));
        const bool acc = de <= 0.0f || u01(word) < __expf(-de);
        if (acc) {
            if (lane == 0) S[k] = static_cast<int8_t>(-s);
            const float ds = -2.0f * fs;
            if constexpr (NQ == 0) {
                const int q0 = AP[k], q1 = AP[k + 1];
                if (a.unique) {
                    for (int q = q0 + lane; q < q1; q += 64) F[AI[q]] += AV[q] * ds;
                } else if (lane == 0) {
                    for (int q = q0; q < q1; ++q) F[AI[q]] += AV[q] * ds;
                }
            } else if constexpr (NQ < 0) {
                // the two-barrier kernel's HUBO update, in the same order (terms in incidence
                // order, products over the term with the flipped spin's new value)
                if (lane == 0) {
                    for (int q = IP[k]; q < IP[k + 1]; ++q) {
                        const int tt = II[q];
                        float prod = TC[tt];
                        for (int e = TO[tt]; e < TO[tt + 1]; ++e) prod *= S[TV[e]];
                        for (int e = TO[tt]; e < TO[tt + 1]; ++e) {
                            const int v = TV[e];
                            if (v != k) F[v] += 2.0f * prod * S[v];
                        }
                    }
                }
            } else {
                f4 hd[NR];
                if constexpr (kDemand) load_row(hd, k);
                const f4 (&hr)[NR] = kDemand ? hd : h;
                static_for<(NQ > 0 ? NQ : 0)>([&](auto q) {
                    const int gi = lane + 64 * q;
                    if (gi < ng) {
                        const f4 x = hr[q], y = F4[gi];
                        F4[gi] = f4{fmaf(ds, x.x, y.x), fmaf(ds, x.y, y.y), fmaf(ds, x.z, y.z), fmaf(ds, x.w, y.w)};
                    }
                });
            }
        }
        __syncthreads();  // this site's field and spin writes before the next site's reads
    };
    // site k of phase ph is step tau = ph * n + k of the sweep: word tau & 3 of Philox block tau >> 2;
    // the words are evaluated a chunk of 64 blocks at a time (one per lane)
    const uint32_t tau0 = static_cast<uint32_t>(a.ph) * static_cast<uint32_t>(n);
    constexpr bool kStream = NQ > 0 && !kDemand;
    f4 A[NR], B[NR];
    if constexpr (kStream) {
        load_row(A, 0);
        load_row(B, 1);
    }
    for (int k = 0; k < n;) {
        const uint32_t b0 = (tau0 + static_cast<uint32_t>(k)) >> 2;
        {
            const Philox4 w = philox4x32_10(static_cast<uint32_t>(c), a.sweep_id, b0 + lane, TAG_CHECKER, a.k0, a.k1);
            W[4 * lane] = w.x;
            W[4 * lane + 1] = w.y;
            W[4 * lane + 2] = w.z;
            W[4 * lane + 3] = w.w;
        }
        __syncthreads();
        const int kend = min(n, static_cast<int>(4 * b0 + kChkWords - tau0));
        for (; k < kend; ++k) {
            const uint32_t word = W[tau0 + k - 4 * b0];
            if constexpr (kStream) {
                if ((k & 1) == 0) {
                    site(A, k, word);
                    load_row(A, k + 2);
                } else {
                    site(B, k, word);
                    load_row(B, k + 2);
                }
            } else {
                site(A, k, word);
            }
        }
        // (the last site's barrier orders this chunk's word reads before the next chunk's writes)
    }
    for (int i = lane; i < n; i += 64) {
        a.gS[off + i] = S[i];
        a.gF[off + i] = F[i];
    }
}

// Pilot magnetisation after a measured sweep (the checkerboard kernel's measurement, same
// arithmetic): one workgroup per replica, m_r = sum of the replica's spins / (n P), accumulated
// as m^2 and |m|.
__global__ void __launch_bounds__(256) chk_mag_kernel(const int8_t *gS, int P, int n, double *m2, double *mabs) {
    __shared__ double ered[kMaxSlices];
    const int r = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int t = wave; t < P; t += 4) {
        const int8_t *sc = gS + (static_cast<int64_t>(r) * P + t) * n;
        int sum = 0;
        for (int i = lane; i < n; i += 64) sum += sc[i];
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (lane == 0) ered[t] = static_cast<double>(sum);
    }
    __syncthreads();
    if (tid == 0) {
        double tot = 0.0;
        for (int t = 0; t < P; ++t) tot += ered[t];
        const double m = tot / (static_cast<double>(n) * P);
        m2[r] += m * m;
        mabs[r] += fabs(m);
    }
}

// Random start state of the wavefront path for dense problems (group_init's fresh branch, same
// arithmetic): spins from Philox (chain, i, TAG_INIT), fields f_ci = h_i + sum_j J_ji s_cj summed in
// FP64 in j order and rounded to FP32. A workgroup takes 256 sites of kChkInitChains chains, so
// each J column it reads serves all of them.
constexpr int kChkInitChains = 16;

__global__ void __launch_bounds__(256) chk_init_kernel(ProblemView pv, int chains, uint32_t k0, uint32_t k1,
                                                       int8_t *gS, float *gF) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = pv.n, tid = threadIdx.x;
    const int c0 = blockIdx.y * kChkInitChains, nc = min(kChkInitChains, chains - c0);
    int8_t *S = reinterpret_cast<int8_t *>(smem);  // [kChkInitChains][n]
    for (int e = tid; e < nc * n; e += 256) {
        const int q = e / n, i = e - q * n;
        const Philox4 r = philox4x32_10(static_cast<uint32_t>(c0 + q), static_cast<uint32_t>(i), 0u, TAG_INIT, k0, k1);
        const int8_t v = (r.x & 1u) ? 1 : -1;
        S[e] = v;
        if (blockIdx.x == 0) gS[static_cast<int64_t>(c0) * n + e] = v;
    }
    __syncthreads();
    const int i = blockIdx.x * 256 + tid;
    if (i >= n) return;
    double acc[kChkInitChains];
#pragma unroll
    for (int q = 0; q < kChkInitChains; ++q) acc[q] = pv.h32[i];
    for (int j = 0; j < n; ++j) {
        const double jv = static_cast<double>(pv.J32[static_cast<int64_t>(j) * n + i]);
#pragma unroll
        for (int q = 0; q < kChkInitChains; ++q)
            if (q < nc) acc[q] += jv * S[q * n + j];
    }
#pragma unroll
    for (int q = 0; q < kChkInitChains; ++q)
        if (q < nc) gF[static_cast<int64_t>(c0 + q) * n + i] = static_cast<float>(acc[q]);
}

void chk_init(const qa_problem *p, int64_t chains, uint32_t k0, uint32_t k1, int8_t *gS, float *gF, hipStream_t s) {
    const int n = static_cast<int>(p->n);
    const size_t shmem = static_cast<size_t>(kChkInitChains) * n;
    ensure_dynamic_lds(reinterpret_cast<const void *>(chk_init_kernel), shmem);
    hipLaunchKernelGGL(chk_init_kernel, dim3(static_cast<unsigned>((n + 255) / 256),
                                             static_cast<unsigned>((chains + kChkInitChains - 1) / kChkInitChains)),
                       dim3(256), shmem, s, view_of(p), static_cast<int>(chains), k0, k1, gS, gF);
    QA_HIP(hipGetLastError());
}

// Step tail of the wavefront path (group_tail's tail_mode 0 with the same arithmetic): slice
// energies E_c = c + 1/2 sum_i s_i (h_i + f_i), one wavefront per chain (lane-strided sums, the
// half taken per lane, then the xor-shuffle reduction), ...
__global__ void __launch_bounds__(64) chk_energy_kernel(ProblemView pv, const int8_t *gS, const float *gF,
                                                        double *chain_e) {
    const int c = blockIdx.x, lane = threadIdx.x, n = pv.n;
    const int8_t *sc = gS + static_cast<int64_t>(c) * n;
    const float *fc = gF + static_cast<int64_t>(c) * n;
    double acc = 0.0;
    if (pv.kind != QA_KIND_HUBO) {
        for (int i = lane; i < n; i += 64)
            acc += static_cast<double>(sc[i]) * (static_cast<double>(pv.h32[i]) + static_cast<double>(fc[i]));
        acc *= 0.5;
    } else {  // direct sum, as group_tail
        for (int i = lane; i < n; i += 64) acc += static_cast<double>(pv.h32[i]) * sc[i];
        for (int t = lane; t < pv.nterms; t += 64) {
            double prod = pv.t_coef32[t];
            for (int q = pv.t_off[t]; q < pv.t_off[t + 1]; ++q) prod *= sc[pv.t_vars[q]];
            acc += prod;
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) chain_e[c] = pv.c + acc;
}

// ... then per replica group (rg replicas) the chain energies summed in chain order into the trace
// partial and the group's first strict minimum kept with its state.
__global__ void __launch_bounds__(256) chk_tail_kernel(const double *chain_e, const int8_t *gS, int R, int P, int n,
                                                       int rg, int ngroups, int step, double *trace_part,
                                                       double *best_e, int *best_step, int *best_chain,
                                                       int8_t *best_state) {
    __shared__ int s_arg;
    const int g = blockIdx.x, tid = threadIdx.x;
    const int r0 = g * rg, G = min(rg, R - r0) * P;
    if (tid == 0) {
        double sum = 0.0, mn = best_e[g];
        int arg = -1;
        for (int c = 0; c < G; ++c) {
            const double e = chain_e[static_cast<int64_t>(r0) * P + c];
            sum += e;
            if (e < mn) {
                mn = e;
                arg = c;
            }
        }
        trace_part[static_cast<int64_t>(step) * ngroups + g] = sum;
        if (arg >= 0) {
            best_e[g] = mn;
            best_step[g] = step;
            best_chain[g] = r0 * P + arg;
        }
        s_arg = arg;
    }
    __syncthreads();
    const int arg = s_arg;
    if (arg >= 0)
        for (int i = tid; i < n; i += 256)
            best_state[static_cast<int64_t>(g) * n + i] = gS[(static_cast<int64_t>(r0) * P + arg) * n + i];
}

// Problems whose checkerboard sweeps run in chk_wave_kernel: dense up to this size (the J rows are
// loaded on demand past 2048 sites), sparse and HUBO problems whose wave state fits LDS. For the
// pilot scan (64x more chains) sparse and HUBO coupling lists are read from global memory (L2)
// instead of staged per wave, so that a wave's LDS stays small and many waves share a CU.
constexpr int kChkWaveMaxN = 8192;
// Dense J rows are loaded on demand (chk_wave_kernel<NQ, true>: only for accepted flips) rather
// than streamed ahead of every site: the other waves in flight hide each load's latency and only
// the accepted fraction of rows is read. Measured on SK at the C5 shape: pilot N=2048 243-354 ms
// vs 503-597 ms streamed, N=512 13-17 vs 22-24 ms; a whole run_chi at N=2048 1.385 s vs 1.784 s.
// QANNEAL_CHI_ROWS=stream selects the streamed rows (up to 2048 sites).
bool chk_rows_on_demand() {
    const char *env = std::getenv("QANNEAL_CHI_ROWS");
    return !(env && env[0] == 's');
}

// LDS bytes of a wave's state plus, when staged, its coupling lists; 0 when lists are not staged.
size_t chk_list_bytes_of(const qa_problem *p) {
    return chk_list_bytes(p->kind, static_cast<int>(p->n), static_cast<int>(2 * p->nedges), static_cast<int>(p->nterms),
                          static_cast<int>(p->t_vars.count));
}

bool chk_lists_staged(const qa_problem *p, bool pilot) {
    const int n = static_cast<int>(p->n), rn = (n + 3) / 4 * 4;
    return !pilot && p->kind != QA_KIND_DENSE && chk_lds_bytes(n, rn) + chk_list_bytes_of(p) <= 160 * 1024;
}

bool chk_wave_applies(const qa_problem *p, bool pilot) {
    const char *env = std::getenv("QANNEAL_CHI_WAVE");
    if (env && env[0] == '0') return false;
    const int n = static_cast<int>(p->n), rn = (n + 3) / 4 * 4;
    if (p->kind == QA_KIND_DENSE) return n <= kChkWaveMaxN;
    return chk_lds_bytes(n, rn) <= 160 * 1024;
}

// Launches the phases of one checkerboard sweep (sweep_id) of every chain; `demand`: dense rows
// loaded only for accepted flips (chk_wave_kernel<NQ, true>); `pilot`: lists read from global memory.
void chk_sweep(const qa_problem *p, ChkArgs &ca, int64_t chains, int P, hipStream_t s, bool demand, bool pilot) {
    const int n = static_cast<int>(p->n), rn = (n + 3) / 4 * 4, ngr = rn / 4;
    const int nnz = static_cast<int>(2 * p->nedges);
    const int kind = p->kind;
    if (ngr > 512) demand = true;  // rows of more than 2048 floats are never streamed
    QA_REQUIRE(kind != QA_KIND_DENSE || ngr <= 2048, "chk_wave_kernel: row too long");
    const void *kern = kind == QA_KIND_SPARSE ? reinterpret_cast<const void *>(chk_wave_kernel<0>)
                     : kind == QA_KIND_HUBO   ? reinterpret_cast<const void *>(chk_wave_kernel<-1>)
                     : demand ? (ngr <= 64 ? reinterpret_cast<const void *>(chk_wave_kernel<1, true>)
                                 : ngr <= 128 ? reinterpret_cast<const void *>(chk_wave_kernel<2, true>)
                                 : ngr <= 256 ? reinterpret_cast<const void *>(chk_wave_kernel<4, true>)
                                 : ngr <= 512 ? reinterpret_cast<const void *>(chk_wave_kernel<8, true>)
                                 : ngr <= 1024 ? reinterpret_cast<const void *>(chk_wave_kernel<16, true>)
                                               : reinterpret_cast<const void *>(chk_wave_kernel<32, true>))
                     : ngr <= 64 ? reinterpret_cast<const void *>(chk_wave_kernel<1>)
                     : ngr <= 128 ? reinterpret_cast<const void *>(chk_wave_kernel<2>)
                     : ngr <= 256 ? reinterpret_cast<const void *>(chk_wave_kernel<4>)
                                  : reinterpret_cast<const void *>(chk_wave_kernel<8>);
    const bool staged = chk_lists_staged(p, pilot);
    const size_t shmem = chk_lds_bytes(n, rn) + (staged ? chk_list_bytes_of(p) : 0);
    QA_REQUIRE(shmem <= 160 * 1024, "chk_wave_kernel: wave state does not fit LDS");
    ensure_dynamic_lds(kern, shmem);
    const ProblemView pv = view_of(p);
    ca.J4rows = pv.J32p;
    ca.adj_ptr = pv.adj_ptr;
    ca.adj_idx = pv.adj_idx;
    ca.adj_val = pv.adj_val32;
    ca.inc_ptr = pv.inc_ptr;
    ca.inc_idx = pv.inc_idx;
    ca.t_off = pv.t_off;
    ca.t_vars = pv.t_vars;
    ca.t_coef = pv.t_coef32;
    ca.nnz = nnz;
    ca.nterms = static_cast<int>(p->nterms);
    ca.nv = static_cast<int>(p->t_vars.count);
    ca.lists_lds = staged ? 1 : 0;
    ca.unique = p->adj_unique ? 1 : 0;
    ca.P = P;
    ca.n = n;
    ca.rn = rn;
    const int nph = P == 1 ? 1 : ((P & 1) ? 3 : 2);
    for (int ph = 0; ph < nph; ++ph) {
        ca.ph = ph;
        void *params[] = {&ca};
        QA_HIP(hipLaunchKernel(kern, dim3(static_cast<unsigned>(chains)), dim3(64), params, shmem, s));
    }
}
}  // namespace

// sqa_chi pilot scan: `points` x R replicas in one launch; replica rep belongs to point rep / R
// and sweeps (burn + sweeps) checkerboard sweeps on a fresh worldline at that point's J_perp,
// accumulating sum m^2 and sum |m| over the measured sweeps.
void sqa_fp32_pilot(qa_ctx *ctx, const qa_problem *p, int64_t P, int64_t R, int64_t points,
                    const std::vector<double> &bs, const std::vector<double> &jp, int64_t burn, int64_t sweeps,
                    uint64_t seed, std::vector<double> &m2, std::vector<double> &mabs, float *ms) {
    const int64_t n = p->n, NS = points * R;
    const SqaPlan pl = plan_sqa(ctx, p, NS, P);
    hipStream_t s = ctx->stream;
    std::vector<float> hbs(points), hjp(points), hjoin(points);
    std::vector<int> pof(NS);
    for (int64_t q = 0; q < points; ++q) {
        hbs[q] = static_cast<float>(bs[q]);
        hjp[q] = static_cast<float>(jp[q]);
        hjoin[q] = static_cast<float>(1.0 - std::exp(-2.0 * jp[q]));
    }
    for (int64_t k = 0; k < NS; ++k) pof[k] = static_cast<int>(k / R);
    const bool wave = chk_wave_applies(p, true);
    DevBuf<float> dbs(points), djp(points), djoin(points), gF(pl.use_lds && !wave ? 1 : NS * P * n);
    DevBuf<int> dpof(NS);
    DevBuf<int8_t> gS(NS * P * n);
    DevBuf<double> dm2(NS), dma(NS);
    DevBuf<unsigned long long> accd(pl.ngroups);
    dbs.upload(hbs.data(), points, s);
    djp.upload(hjp.data(), points, s);
    djoin.upload(hjoin.data(), points, s);
    dpof.upload(pof.data(), NS, s);
    QA_HIP(hipMemsetAsync(dm2.ptr, 0, NS * sizeof(double), s));
    QA_HIP(hipMemsetAsync(dma.ptr, 0, NS * sizeof(double), s));
    SqaArgs a{};
    a.pv = view_of(p);
    a.P = static_cast<int>(P);
    a.R = static_cast<int>(NS);
    a.rg = static_cast<int>(pl.rg);
    a.steps = 1;
    a.spb = static_cast<int>(burn + sweeps);
    a.bs = dbs.ptr;
    a.jp = djp.ptr;
    a.join = djoin.ptr;
    a.k0 = static_cast<uint32_t>(seed);
    a.k1 = static_cast<uint32_t>(seed >> 32);
    a.gS = gS.ptr;
    a.gF = gF.ptr;
    a.accepted = accd.ptr;
    a.param_of_rep = dpof.ptr;
    a.tail_mode = 2;
    a.sweep_kind = 1;
    a.mag_burn = static_cast<int>(burn);
    a.mag_m2 = dm2.ptr;
    a.mag_abs = dma.ptr;
    QA_HIP(hipEventRecord(ctx->ev0, s));
    if (!wave) {
        launch_sqa(ctx, pl, a);
    } else {
        // random start state and its fields, then each sweep's phases in chk_wave_kernel and,
        // from sweep `burn` on, the magnetisation sample
        if (p->kind == QA_KIND_DENSE) {
            chk_init(p, NS * P, a.k0, a.k1, gS.ptr, gF.ptr, s);
        } else {
            SqaArgs init = a;
            init.spb = 0;
            init.save_fields = 1;
            init.mag_m2 = nullptr;
            init.mag_abs = nullptr;
            launch_sqa(ctx, pl, init);
        }
        ChkArgs ca{};
        ca.gS = gS.ptr;
        ca.gF = gF.ptr;
        ca.bsv = dbs.ptr;
        ca.jpv = djp.ptr;
        ca.pof = dpof.ptr;
        ca.k0 = a.k0;
        ca.k1 = a.k1;
        const bool demand = chk_rows_on_demand();
        for (int64_t sw = 0; sw < burn + sweeps; ++sw) {
            ca.sweep_id = static_cast<uint32_t>(sw);
            chk_sweep(p, ca, NS * P, static_cast<int>(P), s, demand, true);
            if (sw >= burn) {
                hipLaunchKernelGGL(chk_mag_kernel, dim3(static_cast<unsigned>(NS)), dim3(256), 0, s, gS.ptr,
                                   static_cast<int>(P), static_cast<int>(n), dm2.ptr, dma.ptr);
                QA_HIP(hipGetLastError());
            }
        }
    }
    QA_HIP(hipEventRecord(ctx->ev1, s));
    m2.resize(NS);
    mabs.resize(NS);
    dm2.download(m2.data(), NS, s);
    dma.download(mabs.data(), NS, s);
    QA_HIP(hipStreamSynchronize(s));
    QA_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
}

// sqa_chi production anneal: R replicas, checkerboard sweeps along (beta_k, Gamma_k).

void sqa_fp32_chi_production(qa_ctx *ctx, const qa_problem *p, int64_t P, int64_t R, int64_t steps,
                             const double *betas, const double *gammas, int64_t sweeps, uint64_t seed,
                             int8_t *best_state, double *best_energy, double *trace, float *ms) {
    const int64_t n = p->n;
    const SqaPlan pl = plan_sqa(ctx, p, R, P);
    const int64_t ngroups = pl.ngroups;
    hipStream_t s = ctx->stream;
    std::vector<float> hbs(steps), hjp(steps), hjoin(steps);
    for (int64_t k = 0; k < steps; ++k) {
        const double jp = trotter_coupling(betas[k], gammas[k], P);
        hbs[k] = static_cast<float>(betas[k] / static_cast<double>(P));
        hjp[k] = static_cast<float>(jp);
        hjoin[k] = static_cast<float>(1.0 - std::exp(-2.0 * jp));
    }
    const bool wave = chk_wave_applies(p, false);
    DevBuf<float> dbs(steps), djp(steps), djoin(steps), gF(pl.use_lds && !wave ? 1 : R * P * n);
    DevBuf<int8_t> gS(R * P * n), bstate(ngroups * n);
    DevBuf<double> tpart(steps * ngroups), be(ngroups), ex(1);
    DevBuf<int> bstep(ngroups), bchain(ngroups);
    DevBuf<unsigned long long> accd(ngroups);
    dbs.upload(hbs.data(), steps, s);
    djp.upload(hjp.data(), steps, s);
    djoin.upload(hjoin.data(), steps, s);
    std::vector<double> inf(ngroups, std::numeric_limits<double>::infinity());
    be.upload(inf.data(), ngroups, s);
    QA_HIP(hipMemsetAsync(bchain.ptr, 0xff, ngroups * sizeof(int), s));
    SqaArgs a{};
    a.pv = view_of(p);
    a.P = static_cast<int>(P);
    a.R = static_cast<int>(R);
    a.rg = static_cast<int>(pl.rg);
    a.steps = static_cast<int>(steps);
    a.spb = static_cast<int>(sweeps);
    a.bs = dbs.ptr;
    a.jp = djp.ptr;
    a.join = djoin.ptr;
    a.k0 = static_cast<uint32_t>(seed);
    a.k1 = static_cast<uint32_t>(seed >> 32);
    a.gS = gS.ptr;
    a.gF = gF.ptr;
    a.trace_part = tpart.ptr;
    a.best_e = be.ptr;
    a.best_step = bstep.ptr;
    a.best_chain = bchain.ptr;
    a.best_state = bstate.ptr;
    a.accepted = accd.ptr;
    a.sweep_kind = 1;
    QA_HIP(hipEventRecord(ctx->ev0, s));
    if (!wave) {
        launch_sqa(ctx, pl, a);
    } else {
        // random start state and its fields (no sweeps, no tail), then per schedule step the
        // phases of every sweep in chk_wave_kernel and the step tail in the two-barrier kernel
        if (p->kind == QA_KIND_DENSE) {
            chk_init(p, R * P, a.k0, a.k1, gS.ptr, gF.ptr, s);
        } else {
            SqaArgs init = a;
            init.steps = 1;
            init.spb = 0;
            init.tail_mode = 2;
            init.save_fields = 1;
            launch_sqa(ctx, pl, init);
        }
        DevBuf<double> ce(R * P);  // chain energies of a step tail
        ChkArgs ca{};
        ca.gS = gS.ptr;
        ca.gF = gF.ptr;
        ca.bsv = dbs.ptr;
        ca.jpv = djp.ptr;
        ca.k0 = a.k0;
        ca.k1 = a.k1;
        for (int64_t k = 0; k < steps; ++k) {
            ca.pidx = static_cast<int>(k);
            for (int64_t sw = 0; sw < sweeps; ++sw) {
                ca.sweep_id = static_cast<uint32_t>(k * sweeps + sw);
                chk_sweep(p, ca, R * P, static_cast<int>(P), s, chk_rows_on_demand(), false);
            }
            hipLaunchKernelGGL(chk_energy_kernel, dim3(static_cast<unsigned>(R * P)), dim3(64), 0, s, a.pv, gS.ptr,
                               gF.ptr, ce.ptr);
            hipLaunchKernelGGL(chk_tail_kernel, dim3(static_cast<unsigned>(ngroups)), dim3(256), 0, s, ce.ptr, gS.ptr,
                               static_cast<int>(R), static_cast<int>(P), static_cast<int>(n),
                               static_cast<int>(pl.rg), static_cast<int>(ngroups), static_cast<int>(k), tpart.ptr,
                               be.ptr, bstep.ptr, bchain.ptr, bstate.ptr);
        }
        QA_HIP(hipGetLastError());
    }
    QA_HIP(hipEventRecord(ctx->ev1, s));
    std::vector<double> hpart(steps * ngroups), hbe(ngroups);
    std::vector<int> hstep(ngroups), hchain(ngroups);
    tpart.download(hpart.data(), hpart.size(), s);
    be.download(hbe.data(), ngroups, s);
    bstep.download(hstep.data(), ngroups, s);
    bchain.download(hchain.data(), ngroups, s);
    QA_HIP(hipStreamSynchronize(s));
    QA_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    for (int64_t k = 0; k < steps; ++k) {
        double sum = 0.0;
        for (int64_t gi = 0; gi < ngroups; ++gi) sum += hpart[k * ngroups + gi];
        trace[k] = sum / static_cast<double>(R * P);
    }
    int64_t bg = -1;
    for (int64_t gi = 0; gi < ngroups; ++gi) {
        if (hchain[gi] < 0) continue;
        if (bg < 0 || hbe[gi] < hbe[bg] || (hbe[gi] == hbe[bg] && hstep[gi] < hstep[bg])) bg = gi;
    }
    QA_REQUIRE(bg >= 0, "no finite slice energy was produced");
    QA_HIP(hipMemcpyAsync(best_state, bstate.ptr + bg * n, n, hipMemcpyDeviceToHost, s));
    device_energies(p, bstate.ptr + bg * n, 1, ex.ptr, s);
    ex.download(best_energy, 1, s);
    QA_HIP(hipStreamSynchronize(s));
}

// ---- batched SQA (many instances, one persistent launch per sweep variant) --------------------
namespace {

constexpr size_t kBatchLdsCap = 160 * 1024 - 64;  // room for the batch kernel's static LDS

// Sweep variant a plan runs as: pipelined with 1/2/4/8 row groups per lane, or two-barrier with
// the state in LDS (0) or global memory (-1).
int batch_variant(const qa_problem *p, const SqaPlan &pl, const qa_batch_params *prm) {
    const bool pipe = pl.pipe && prm->worldline_sweeps == 0 && prm->cluster_sweeps == 0 && p->kind == QA_KIND_DENSE;
    if (pipe) {
        const int ngr = pl.ldf / 4;
        return ngr <= 64 ? 1 : ngr <= 128 ? 2 : ngr <= 256 ? 4 : 8;
    }
    return pl.use_lds ? 0 : -1;
}

const void *batch_kernel_of(int v) {
    switch (v) {
        case 1: return reinterpret_cast<const void *>(sqa_batch_kernel<1>);
        case 2: return reinterpret_cast<const void *>(sqa_batch_kernel<2>);
        case 4: return reinterpret_cast<const void *>(sqa_batch_kernel<4>);
        case 8: return reinterpret_cast<const void *>(sqa_batch_kernel<8>);
        case 0: return reinterpret_cast<const void *>(sqa_batch_kernel<0>);
        default: return reinterpret_cast<const void *>(sqa_batch_kernel<-1>);
    }
}

// Autotuner: replicas per group for (model, n rounded up to 64, P). Each candidate group size is
// timed on a trial that fills every CU with one group of this problem for two sweeps; the fastest
// attempts/s wins and is cached in the context.
int64_t tuned_rg(qa_ctx *ctx, const qa_problem *p, const qa_batch_params *prm) {
    const int64_t P = prm->trotter_slices;
    const auto key = std::make_tuple(p->kind, (p->n + 63) / 64 * 64, P);
    const auto hit = ctx->tuned_rg.find(key);
    if (hit != ctx->tuned_rg.end()) return hit->second;
    hipStream_t s = ctx->stream;
    const int64_t n = p->n;
    int64_t best_rg = 1;
    double best_rate = -1.0;
    const float bs = static_cast<float>(prm->betas[0] / static_cast<double>(P));
    const double jpd = trotter_coupling(prm->betas[0], prm->gammas[0], P);
    const float jp = static_cast<float>(jpd), join = static_cast<float>(1.0 - std::exp(-2.0 * jpd));
    DevBuf<float> dbs(1), djp(1), djoin(1);
    dbs.upload(&bs, 1, s);
    djp.upload(&jp, 1, s);
    djoin.upload(&join, 1, s);
    for (int64_t rg = 1; rg * P <= kMaxChains; rg *= 2) {
        const SqaPlan pl = plan_sqa(ctx, p, rg * ctx->num_cus, P, rg, kBatchLdsCap);
        if (pl.rg != rg) continue;
        const int64_t R = rg * ctx->num_cus, ng = pl.ngroups;
        DevBuf<int8_t> gS(R * P * n), bstate(ng * n);
        DevBuf<float> gF(pl.use_lds ? 1 : R * P * n);
        DevBuf<double> tpart(ng), be(ng);
        DevBuf<int> bstep(ng), bchain(ng);
        DevBuf<unsigned long long> accd(ng);
        std::vector<double> inf(ng, std::numeric_limits<double>::infinity());
        be.upload(inf.data(), ng, s);
        SqaArgs a{};
        a.pv = view_of(p);
        a.P = static_cast<int>(P);
        a.R = static_cast<int>(R);
        a.rg = static_cast<int>(rg);
        a.steps = 1;
        a.spb = 2;
        a.wl = static_cast<int>(std::min<int64_t>(prm->worldline_sweeps, 1));
        a.cl = static_cast<int>(std::min<int64_t>(prm->cluster_sweeps, 1));
        a.bs = dbs.ptr;
        a.jp = djp.ptr;
        a.join = djoin.ptr;
        a.gS = gS.ptr;
        a.gF = gF.ptr;
        a.trace_part = tpart.ptr;
        a.best_e = be.ptr;
        a.best_step = bstep.ptr;
        a.best_chain = bchain.ptr;
        a.best_state = bstate.ptr;
        a.accepted = accd.ptr;
        QA_HIP(hipEventRecord(ctx->ev0, s));
        launch_sqa(ctx, pl, a);
        QA_HIP(hipEventRecord(ctx->ev1, s));
        QA_HIP(hipStreamSynchronize(s));
        float ms = 0;
        QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        const double rate = static_cast<double>(R * P * n * (a.spb + a.wl + a.cl)) / std::max(1e-3, double(ms));
        if (rate > best_rate) {
            best_rate = rate;
            best_rg = rg;
        }
    }
    ctx->tuned_rg[key] = best_rg;
    return best_rg;
}

}  // namespace

void sqa_batch(qa_ctx *ctx, int64_t count, const qa_problem *const *probs, const qa_batch_params *prm,
               qa_batch_result *res) {
    hipStream_t s = ctx->stream;
    const int64_t P = prm->trotter_slices, steps = prm->steps;
    const int64_t per_step = prm->sweeps_per_beta + prm->worldline_sweeps + prm->cluster_sweeps;
    // shared schedule
    std::vector<float> hbs(steps), hjp(steps), hjoin(steps);
    for (int64_t k = 0; k < steps; ++k) {
        const double jp = trotter_coupling(prm->betas[k], prm->gammas[k], P);
        hbs[k] = static_cast<float>(prm->betas[k] / static_cast<double>(P));
        hjp[k] = static_cast<float>(jp);
        hjoin[k] = static_cast<float>(1.0 - std::exp(-2.0 * jp));
    }
    DevBuf<float> dbs(steps), djp(steps), djoin(steps);
    dbs.upload(hbs.data(), steps, s);
    djp.upload(hjp.data(), steps, s);
    djoin.upload(hjoin.data(), steps, s);

    // plans and arena offsets
    struct Inst {
        SqaPlan pl;
        int variant;
        int64_t R, n, oS, oF, oB, oT, oG;  // offsets: spins, fields, best states, trace, per-group
    };
    std::vector<Inst> in(count);
    int64_t nS = 0, nF = 0, nB = 0, nT = 0, nG = 0, attempts = 0;
    for (int64_t k = 0; k < count; ++k) {
        const qa_problem *p = probs[k];
        Inst &x = in[k];
        x.R = prm->reads[k];
        x.n = p->n;
        const int64_t rg = prm->autotune ? tuned_rg(ctx, p, prm) : 0;
        x.pl = plan_sqa(ctx, p, x.R, P, rg, kBatchLdsCap);
        x.variant = batch_variant(p, x.pl, prm);
        x.oS = nS;
        x.oF = nF;
        x.oB = nB;
        x.oT = nT;
        x.oG = nG;
        nS += x.R * P * x.n;
        if (!x.pl.use_lds) nF += x.R * P * x.n;
        nB += x.pl.ngroups * x.n;
        nT += steps * x.pl.ngroups;
        nG += x.pl.ngroups;
        attempts += steps * per_step * x.R * P * x.n;
    }
    DevBuf<int8_t> gS(nS), bstate(nB);
    DevBuf<float> gF(std::max<int64_t>(nF, 1));
    DevBuf<double> tpart(nT), be(nG);
    DevBuf<int> bstep(nG), bchain(nG);
    DevBuf<unsigned long long> accd(nG);
    std::vector<double> inf(nG, std::numeric_limits<double>::infinity());
    std::vector<int> minus1(nG, -1);
    be.upload(inf.data(), nG, s);
    bchain.upload(minus1.data(), nG, s);

    // per-instance arguments and work items per variant, largest first
    std::vector<SqaArgs> args(count);
    std::map<int, std::vector<std::pair<double, BatchItem>>> work;
    std::map<int, size_t> shmem;
    for (int64_t k = 0; k < count; ++k) {
        const Inst &x = in[k];
        SqaArgs &a = args[k];
        a = SqaArgs{};
        a.pv = view_of(probs[k]);
        a.P = static_cast<int>(P);
        a.R = static_cast<int>(x.R);
        a.rg = static_cast<int>(x.pl.rg);
        a.steps = static_cast<int>(steps);
        a.spb = static_cast<int>(prm->sweeps_per_beta);
        a.wl = static_cast<int>(prm->worldline_sweeps);
        a.cl = static_cast<int>(prm->cluster_sweeps);
        a.bs = dbs.ptr;
        a.jp = djp.ptr;
        a.join = djoin.ptr;
        a.k0 = static_cast<uint32_t>(prm->seeds[k]);
        a.k1 = static_cast<uint32_t>(prm->seeds[k] >> 32);
        a.gS = gS.ptr + x.oS;
        a.gF = gF.ptr + x.oF;
        a.trace_part = tpart.ptr + x.oT;
        a.best_e = be.ptr + x.oG;
        a.best_step = bstep.ptr + x.oG;
        a.best_chain = bchain.ptr + x.oG;
        a.best_state = bstate.ptr + x.oB;
        a.accepted = accd.ptr + x.oG;
        apply_plan(a, x.pl);
        a.s_lds = 0;  // the batch kernel's global-state body keeps the spins in global memory
        a.la = 0;
        const double cost = static_cast<double>(x.pl.rg * P * x.n) * (probs[k]->kind == QA_KIND_DENSE ? x.n : 16);
        for (int64_t gi = 0; gi < x.pl.ngroups; ++gi)
            work[x.variant].push_back({cost, BatchItem{static_cast<int>(k), static_cast<int>(gi)}});
        shmem[x.variant] = std::max(shmem[x.variant], x.pl.shmem);
    }
    DevBuf<SqaArgs> dargs(count);
    dargs.upload(args.data(), count, s);
    std::vector<BatchItem> items;
    struct Launch {
        const void *kern;
        int64_t first, n;
        size_t shmem;
        unsigned grid;
    };
    std::vector<Launch> launches;
    for (auto &kv : work) {
        auto &v = kv.second;
        std::stable_sort(v.begin(), v.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
        Launch L{batch_kernel_of(kv.first), static_cast<int64_t>(items.size()), static_cast<int64_t>(v.size()),
                 shmem[kv.first], 0};
        for (auto &w : v) items.push_back(w.second);
        const int64_t per_cu = std::max<int64_t>(1, static_cast<int64_t>((160 * 1024) / (L.shmem + 64)));
        L.grid = static_cast<unsigned>(std::min<int64_t>(L.n, per_cu * ctx->num_cus));
        ensure_dynamic_lds(L.kern, L.shmem);
        launches.push_back(L);
    }
    DevBuf<BatchItem> ditems(items.size());
    ditems.upload(items.data(), items.size(), s);
    // [0, L): item counters of the L launches; [L]: items completed over all launches
    DevBuf<int> counters(launches.size() + 1);

    auto enqueue = [&]() {
        hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counters.ptr,
                           static_cast<int>(launches.size() + 1));
        QA_HIP(hipGetLastError());
        for (size_t li = 0; li < launches.size(); ++li) {
            const SqaArgs *pa = dargs.ptr;
            const BatchItem *pi = ditems.ptr + launches[li].first;
            int ni = static_cast<int>(launches[li].n);
            int *pc = counters.ptr + li;
            int *pd = counters.ptr + launches.size();
            void *params[] = {&pa, &pi, &ni, &pc, &pd};
            QA_HIP(hipLaunchKernel(launches[li].kern, dim3(launches[li].grid), dim3(kThreads), params,
                                   launches[li].shmem, s));
        }
    };
    QA_HIP(hipStreamSynchronize(s));
    QA_HIP(hipEventRecord(ctx->ev0, s));
    if (prm->use_graph) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        QA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue();
        QA_HIP(hipStreamEndCapture(s, &graph));
        QA_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        QA_HIP(hipGraphLaunch(exec, s));
        QA_HIP(hipStreamSynchronize(s));
        QA_HIP(hipGraphExecDestroy(exec));
        QA_HIP(hipGraphDestroy(graph));
    } else {
        enqueue();
    }
    QA_HIP(hipEventRecord(ctx->ev1, s));
    int done_items = 0;
    QA_HIP(hipMemcpyAsync(&done_items, counters.ptr + launches.size(), sizeof(int), hipMemcpyDeviceToHost, s));
    QA_HIP(hipStreamSynchronize(s));
    if (done_items != static_cast<int>(items.size()))
        throw Error{QA_EDEVICE, "qa_sqa_batch_run: the batched launch completed " + std::to_string(done_items) +
                                    " of " + std::to_string(items.size()) + " work items"};

    // per instance: best group (first strict minimum in (step, replica, slice) order), FP64 energy
    std::vector<double> hbe(nG);
    std::vector<int> hstep(nG), hchain(nG);
    be.download(hbe.data(), nG, s);
    bstep.download(hstep.data(), nG, s);
    bchain.download(hchain.data(), nG, s);
    QA_HIP(hipStreamSynchronize(s));
    float ms = 0;
    QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    DevBuf<double> ex(count);
    int64_t off = 0;
    for (int64_t k = 0; k < count; ++k) {
        const Inst &x = in[k];
        int64_t bg = -1;
        for (int64_t gi = 0; gi < x.pl.ngroups; ++gi) {
            const int64_t q = x.oG + gi;
            if (hchain[q] < 0) continue;
            if (bg < 0 || hbe[q] < hbe[x.oG + bg] || (hbe[q] == hbe[x.oG + bg] && hstep[q] < hstep[x.oG + bg])) bg = gi;
        }
        QA_REQUIRE(bg >= 0, "no finite slice energy was produced");
        const int8_t *src = bstate.ptr + x.oB + bg * x.n;
        QA_HIP(hipMemcpyAsync(res->best_states + off, src, x.n, hipMemcpyDeviceToHost, s));
        device_energies(probs[k], src, 1, ex.ptr + k, s);
        off += x.n;
    }
    ex.download(res->best_energies, count, s);
    QA_HIP(hipStreamSynchronize(s));
    if (res->kernel_ms) *res->kernel_ms = ms;
    if (res->attempts) *res->attempts = attempts;
}

}  // namespace qa
