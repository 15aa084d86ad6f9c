// FP64 validation mode: the device replays the reference CPU trajectory.
//
// Arithmetic follows the reference operation by operation (FP64 fields, the same summation
// orders, de = fma(beta/P, dE_classical, dE_trotter) where the reference build contracts), and
// the random stream is the reference's: std::mt19937_64 outputs drawn on the host, consumed on
// the device exactly where the reference consumes them (a uniform only when dE > 0, slice picks
// through Lemire's rejection). Shuffles of the visit order are drawn on the host with
// std::shuffle on the same engine, so libstdc++ <random> defines every draw.
//
// Residual difference: exp() on the device (ROCm device libm) and in glibc may differ in the
// last ulp; a decision flips only if the uniform falls inside that ulp (~1e-16 per uphill move).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "qa_internal.hpp"
#include "qa_rng.hpp"

namespace qa {
namespace {

// ---- FP64 Hamiltonian pieces on the device (reference order) ------------------------------
__device__ double delta_fp64(const ProblemView &pv, const int8_t *s, int f) {
    // dense_ising.cpp:46-62 / sparse_ising.cpp:58-71 / higher_order_ising.cpp:113-136
    double local = pv.h64[f];
    if (pv.kind == QA_KIND_DENSE) {
        const double *row = pv.J64 + static_cast<int64_t>(f) * pv.n;
        for (int j = 0; j < pv.n; ++j)
            if (j != f) local += row[j] * static_cast<double>(s[j]);
    } else if (pv.kind == QA_KIND_SPARSE) {
        for (int q = pv.adj_ptr[f]; q < pv.adj_ptr[f + 1]; ++q)
            local += pv.adj_val64[q] * static_cast<double>(s[pv.adj_idx[q]]);
    } else {
        for (int q = pv.inc_ptr[f]; q < pv.inc_ptr[f + 1]; ++q) {
            const int t = pv.inc_idx[q];
            double prod = pv.t_coef64[t];
            for (int k = pv.t_off[t]; k < pv.t_off[t + 1]; ++k)
                if (pv.t_vars[k] != f) prod *= static_cast<double>(s[pv.t_vars[k]]);
            local += prod;
        }
    }
    return -2.0 * static_cast<double>(s[f]) * local;
}

// fields[i] for one state; i-parallel over the calling threads (dense_ising.cpp:64-83,
// sparse_ising.cpp:73-85, higher_order_ising.cpp:138-160; per element the contributions arrive
// in the same order as in the reference's loops).
__device__ void fields_fp64(const ProblemView &pv, const int8_t *s, double *f, int lane, int lanes) {
    for (int i = lane; i < pv.n; i += lanes) {
        double acc = pv.h64[i];
        if (pv.kind == QA_KIND_DENSE) {
            const double *row = pv.J64 + static_cast<int64_t>(i) * pv.n;
            for (int j = 0; j < pv.n; ++j) acc += row[j] * static_cast<double>(s[j]);
            acc -= row[i] * static_cast<double>(s[i]);
        } else if (pv.kind == QA_KIND_SPARSE) {
            for (int q = pv.adj_ptr[i]; q < pv.adj_ptr[i + 1]; ++q)
                acc += pv.adj_val64[q] * static_cast<double>(s[pv.adj_idx[q]]);
        } else {
            for (int q = pv.inc_ptr[i]; q < pv.inc_ptr[i + 1]; ++q) {
                const int t = pv.inc_idx[q];
                double prod = pv.t_coef64[t];
                for (int k = pv.t_off[t]; k < pv.t_off[t + 1]; ++k) prod *= static_cast<double>(s[pv.t_vars[k]]);
                acc += prod * static_cast<double>(s[i]);
            }
        }
        f[i] = acc;
    }
}

// update_local_fields_after_flip, wave-cooperative (dense_ising.cpp:88-106; sparse and hubo
// update lists may touch one element twice, so they run in list order on lane 0).
__device__ void update_fp64(const ProblemView &pv, double *f, const int8_t *ns, int flip, int8_t old,
                            int lane, int lanes) {
    if (pv.kind == QA_KIND_DENSE) {
        const double ds = -2.0 * static_cast<double>(old);
        const double *row = pv.J64 + static_cast<int64_t>(flip) * pv.n;
        for (int j = lane; j < pv.n; j += lanes) {
            double v = f[j] + row[j] * ds;
            if (j == flip) v -= row[flip] * ds;
            f[j] = v;
        }
    } else if (lane == 0) {
        if (pv.kind == QA_KIND_SPARSE) {
            const double ds = -2.0 * static_cast<double>(old);
            for (int q = pv.adj_ptr[flip]; q < pv.adj_ptr[flip + 1]; ++q)
                f[pv.adj_idx[q]] += pv.adj_val64[q] * ds;
        } else {
            for (int q = pv.inc_ptr[flip]; q < pv.inc_ptr[flip + 1]; ++q) {
                const int t = pv.inc_idx[q];
                double prod = pv.t_coef64[t];
                for (int k = pv.t_off[t]; k < pv.t_off[t + 1]; ++k) prod *= static_cast<double>(ns[pv.t_vars[k]]);
                for (int k = pv.t_off[t]; k < pv.t_off[t + 1]; ++k) {
                    const int v = pv.t_vars[k];
                    if (v != flip) f[v] += 2.0 * prod * static_cast<double>(ns[v]);
                }
            }
        }
    }
}

__device__ double energy_fp64(const ProblemView &pv, const int8_t *s) {
    double e = pv.c;
    for (int i = 0; i < pv.n; ++i) e += pv.h64[i] * static_cast<double>(s[i]);
    if (pv.kind == QA_KIND_DENSE) {
        for (int i = 0; i < pv.n; ++i) {
            const double si = static_cast<double>(s[i]);
            const double *row = pv.J64 + static_cast<int64_t>(i) * pv.n;
            for (int j = i + 1; j < pv.n; ++j) e += row[j] * si * static_cast<double>(s[j]);
        }
    } else if (pv.kind == QA_KIND_SPARSE) {
        for (int q = 0; q < pv.nedges; ++q)
            e += pv.ev64[q] * static_cast<double>(s[pv.ei[q]]) * static_cast<double>(s[pv.ej[q]]);
    } else {
        for (int t = 0; t < pv.nterms; ++t) {
            double prod = pv.t_coef64[t];
            for (int q = pv.t_off[t]; q < pv.t_off[t + 1]; ++q) prod *= static_cast<double>(s[pv.t_vars[q]]);
            e += prod;
        }
    }
    return e;
}

// ---- SA (annealer.cpp:33-75): one thread per read ----------------------------------------
// Chained runs (optional): read_betas[r] replaces betas[st] for read r; init_e[r] is the running
// energy at the start (the reference carries it between calls, parallel_tempering.cpp:73-85);
// final_e[r] and the spins buffer hold the end state; mag[r][st] the magnetization.
__global__ void sa_validation_kernel(ProblemView pv, int reads, int steps, const double *__restrict__ betas,
                                     int sweeps, int8_t *spins, int8_t *best, double *best_e,
                                     double *trace, const double *__restrict__ uni, int64_t ucap,
                                     int64_t *used, int *status, const double *__restrict__ read_betas,
                                     const double *__restrict__ init_e, double *final_e, double *mag,
                                     double *btrace) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= reads) return;
    const int n = pv.n;
    int8_t *s = spins + static_cast<int64_t>(r) * n;
    int8_t *b = best + static_cast<int64_t>(r) * n;
    const double *u = uni + static_cast<int64_t>(r) * ucap;
    double e = init_e ? init_e[r] : energy_fp64(pv, s);
    double be = e;
    for (int i = 0; i < n; ++i) b[i] = s[i];
    if (btrace) btrace[static_cast<int64_t>(r) * (steps + 1)] = be;
    int64_t cur = 0;
    for (int st = 0; st < steps; ++st) {
        const double beta = read_betas ? read_betas[r] : betas[st];
        for (int sw = 0; sw < sweeps; ++sw)
            for (int i = 0; i < n; ++i) {
                const double d = delta_fp64(pv, s, i);
                bool acc = d <= 0.0;
                if (!acc) {
                    if (cur >= ucap) { status[r] = 1; return; }
                    acc = u[cur++] < exp(-beta * d);
                }
                if (acc) {
                    s[i] = static_cast<int8_t>(-s[i]);
                    e += d;
                    if (e < be) {
                        be = e;
                        for (int k = 0; k < n; ++k) b[k] = s[k];
                    }
                }
            }
        trace[static_cast<int64_t>(r) * steps + st] = e;
        if (btrace) btrace[static_cast<int64_t>(r) * (steps + 1) + st + 1] = be;
        if (mag) {
            int sum = 0;
            for (int i = 0; i < n; ++i) sum += s[i];
            mag[static_cast<int64_t>(r) * steps + st] = static_cast<double>(sum) / static_cast<double>(n);
        }
    }
    best_e[r] = be;
    used[r] = cur;
    if (final_e) final_e[r] = e;
}

// ---- SQA sweep phases (sqa_annealer.cpp:176-315), one wave per replica --------------------
enum { PH_SLICE = 0, PH_CLUSTER = 1, PH_WORLDLINE = 2 };

struct SqaPhaseArgs {
    ProblemView pv;
    int P;
    int phase;
    double bs, jp, join;
    int8_t *spins;         // [R][P][n]
    double *fields;        // [R][P][n]
    const int32_t *order;  // [R][n]
    const uint64_t *stream;// [R][cap]
    int64_t cap;
    int64_t *used;         // [R]
    int *status;           // [R]
};

__global__ void __launch_bounds__(64) sqa_phase_kernel(SqaPhaseArgs a) {
    extern __shared__ unsigned char lds[];
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = a.pv.n, P = a.P;
    int8_t *S = a.spins + static_cast<int64_t>(r) * P * n;
    double *F = a.fields + static_cast<int64_t>(r) * P * n;
    const int32_t *ord = a.order + static_cast<int64_t>(r) * n;
    const uint64_t *str = a.stream + static_cast<int64_t>(r) * a.cap;
    char *incl = reinterpret_cast<char *>(lds);                     // [P]
    int *stk = reinterpret_cast<int *>(lds + ((P + 15) / 16) * 16);  // [P]
    int64_t cur = 0;
    bool ok = true;
    auto sl = [&](int t) { return S + static_cast<int64_t>(t) * n; };
    auto fl = [&](int t) { return F + static_cast<int64_t>(t) * n; };

    if (a.phase == PH_SLICE) {
        for (int t = 0; t < P && ok; ++t) {
            int8_t *cs = sl(t);
            double *f = fl(t);
            const int8_t *pv_ = sl(t == 0 ? P - 1 : t - 1);
            const int8_t *nx = sl((t + 1) % P);
            for (int k = 0; k < n; ++k) {
                const int i = ord[k];
                const int8_t s = cs[i];
                const double dcl = -2.0 * static_cast<double>(s) * f[i];
                const double dtr = 2.0 * a.jp * static_cast<double>(s) * static_cast<double>(pv_[i] + nx[i]);
                const double de = __fma_rn(a.bs, dcl, dtr);  // sqa_annealer.cpp:200
                bool acc = de <= 0.0;
                if (!acc) {
                    if (cur >= a.cap) { ok = false; break; }
                    acc = canonical_from_u64(str[cur++]) < exp(-de);
                }
                if (acc) {
                    __syncthreads();
                    if (lane == 0) cs[i] = static_cast<int8_t>(-s);
                    __syncthreads();
                    update_fp64(a.pv, f, cs, i, s, lane, 64);
                    __syncthreads();
                }
            }
        }
    } else if (a.phase == PH_CLUSTER) {
        for (int k = 0; k < n && ok; ++k) {
            const int i = ord[k];
            const uint64_t seed = lemire_pick(str, &cur, a.cap, static_cast<uint64_t>(P), &ok);
            if (!ok) break;
            // Growth along the imaginary-time ring (:219-246): LIFO, prev before next. Every lane
            // runs the same sequence on identical data; lane 0 stores.
            if (lane == 0) for (int t = 0; t < P; ++t) incl[t] = 0;
            __syncthreads();
            int top = 0;
            if (lane == 0) { stk[0] = static_cast<int>(seed); incl[seed] = 1; }
            __syncthreads();
            top = 1;
            const int8_t s0 = sl(static_cast<int>(seed))[i];
            while (top > 0 && ok) {
                const int t = stk[--top];
                const int pvt = t == 0 ? P - 1 : t - 1, nxt = (t + 1) % P;
                bool join_p = false, join_n = false;
                if (!incl[pvt] && sl(pvt)[i] == s0) {
                    if (cur >= a.cap) { ok = false; break; }
                    join_p = canonical_from_u64(str[cur++]) < a.join;
                }
                __syncthreads();
                if (join_p) {
                    if (lane == 0) { incl[pvt] = 1; stk[top] = pvt; }
                    ++top;
                }
                __syncthreads();
                if (!incl[nxt] && sl(nxt)[i] == s0) {
                    if (cur >= a.cap) { ok = false; break; }
                    join_n = canonical_from_u64(str[cur++]) < a.join;
                }
                __syncthreads();
                if (join_n) {
                    if (lane == 0) { incl[nxt] = 1; stk[top] = nxt; }
                    ++top;
                }
                __syncthreads();
            }
            if (!ok) break;
            double dcl = 0.0, dtm = 0.0;
            for (int t = 0; t < P; ++t) {
                if (!incl[t]) continue;
                dcl += -2.0 * static_cast<double>(s0) * fl(t)[i];
                const int pvt = t == 0 ? P - 1 : t - 1, nxt = (t + 1) % P;
                if (!incl[pvt]) dtm += 2.0 * a.jp * static_cast<double>(s0) * static_cast<double>(sl(pvt)[i]);
                if (!incl[nxt]) dtm += 2.0 * a.jp * static_cast<double>(s0) * static_cast<double>(sl(nxt)[i]);
            }
            const double de = __fma_rn(a.bs, dcl, dtm);  // sqa_annealer.cpp:275
            bool acc = de <= 0.0;
            if (!acc) {
                if (cur >= a.cap) { ok = false; break; }
                acc = canonical_from_u64(str[cur++]) < exp(-de);
            }
            if (acc) {
                for (int t = 0; t < P; ++t) {
                    if (!incl[t]) continue;
                    __syncthreads();
                    if (lane == 0) sl(t)[i] = static_cast<int8_t>(-s0);
                    __syncthreads();
                    update_fp64(a.pv, fl(t), sl(t), i, s0, lane, 64);
                    __syncthreads();
                }
            }
        }
    } else {
        for (int k = 0; k < n; ++k) {
            const int i = ord[k];
            double dcl = 0.0;
            for (int t = 0; t < P; ++t) dcl += -2.0 * static_cast<double>(sl(t)[i]) * fl(t)[i];
            const double de = a.bs * dcl;
            bool acc = de <= 0.0;
            if (!acc) {
                if (cur >= a.cap) { ok = false; break; }
                acc = canonical_from_u64(str[cur++]) < exp(-de);
            }
            if (acc) {
                for (int t = 0; t < P; ++t) {
                    const int8_t old = sl(t)[i];
                    __syncthreads();
                    if (lane == 0) sl(t)[i] = static_cast<int8_t>(-old);
                    __syncthreads();
                    update_fp64(a.pv, fl(t), sl(t), i, old, lane, 64);
                    __syncthreads();
                }
            }
        }
    }
    if (lane == 0) {
        a.used[r] = cur;
        if (!ok) a.status[r] = 1;
    }
}

// ---- SQA-PT sweep phases (sqa_parallel_tempering.cpp:208-296): no field cache, the
// classical dE of every attempt is the O(n) delta_energy; one wave per rung.
struct PtPhaseArgs {
    ProblemView pv;
    int P;
    int phase;
    const double *bs, *jp, *join;  // [rungs]
    const int *slot;               // [rungs] state buffer holding rung r
    int8_t *spins;                 // [rungs][P][n] (indexed by slot)
    const int32_t *order;          // [rungs][n]
    const uint64_t *stream;        // [rungs][cap]
    int64_t cap;
    int64_t *used;
    int *status;
};

__global__ void __launch_bounds__(64) pt_phase_kernel(PtPhaseArgs a) {
    extern __shared__ unsigned char lds[];
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = a.pv.n, P = a.P;
    int8_t *S = a.spins + static_cast<int64_t>(a.slot[r]) * P * n;
    const int32_t *ord = a.order + static_cast<int64_t>(r) * n;
    const uint64_t *str = a.stream + static_cast<int64_t>(r) * a.cap;
    const double bs = a.bs[r], jp = a.jp[r], join = a.join[r];
    char *incl = reinterpret_cast<char *>(lds);
    int *stk = reinterpret_cast<int *>(lds + ((P + 15) / 16) * 16);
    int64_t cur = 0;
    bool ok = true;
    auto sl = [&](int t) { return S + static_cast<int64_t>(t) * n; };
    if (a.phase == PH_SLICE) {
        for (int t = 0; t < P && ok; ++t) {
            int8_t *cs = sl(t);
            const int8_t *pv_ = sl(t == 0 ? P - 1 : t - 1), *nx = sl((t + 1) % P);
            for (int k = 0; k < n; ++k) {
                const int i = ord[k];
                const double dcl = delta_fp64(a.pv, cs, i);
                const double dtr = 2.0 * jp * static_cast<double>(cs[i]) * static_cast<double>(pv_[i] + nx[i]);
                const double de = __fma_rn(bs, dcl, dtr);  // sqa_parallel_tempering.cpp:215
                bool acc = de <= 0.0;
                if (!acc) {
                    if (cur >= a.cap) { ok = false; break; }
                    acc = canonical_from_u64(str[cur++]) < exp(-de);
                }
                __syncthreads();
                if (acc && lane == 0) cs[i] = static_cast<int8_t>(-cs[i]);
                __syncthreads();
            }
        }
    } else if (a.phase == PH_CLUSTER) {
        for (int k = 0; k < n && ok; ++k) {
            const int i = ord[k];
            const uint64_t seed = lemire_pick(str, &cur, a.cap, static_cast<uint64_t>(P), &ok);
            if (!ok) break;
            if (lane == 0) {
                for (int t = 0; t < P; ++t) incl[t] = 0;
                stk[0] = static_cast<int>(seed);
                incl[seed] = 1;
            }
            __syncthreads();
            int top = 1;
            const int8_t s0 = sl(static_cast<int>(seed))[i];
            while (top > 0 && ok) {
                const int t = stk[--top];
                const int pvt = t == 0 ? P - 1 : t - 1, nxt = (t + 1) % P;
                bool jp_ = false, jn_ = false;
                if (!incl[pvt] && sl(pvt)[i] == s0) {
                    if (cur >= a.cap) { ok = false; break; }
                    jp_ = canonical_from_u64(str[cur++]) < join;
                }
                __syncthreads();
                if (jp_) {
                    if (lane == 0) { incl[pvt] = 1; stk[top] = pvt; }
                    ++top;
                }
                __syncthreads();
                if (!incl[nxt] && sl(nxt)[i] == s0) {
                    if (cur >= a.cap) { ok = false; break; }
                    jn_ = canonical_from_u64(str[cur++]) < join;
                }
                __syncthreads();
                if (jn_) {
                    if (lane == 0) { incl[nxt] = 1; stk[top] = nxt; }
                    ++top;
                }
                __syncthreads();
            }
            if (!ok) break;
            double dcl = 0.0, dtm = 0.0;
            for (int t = 0; t < P; ++t) {
                if (!incl[t]) continue;
                dcl += delta_fp64(a.pv, sl(t), i);
                const int pvt = t == 0 ? P - 1 : t - 1, nxt = (t + 1) % P;
                const double st = static_cast<double>(sl(t)[i]);
                if (!incl[pvt]) dtm += 2.0 * jp * st * static_cast<double>(sl(pvt)[i]);
                if (!incl[nxt]) dtm += 2.0 * jp * st * static_cast<double>(sl(nxt)[i]);
            }
            const double de = __fma_rn(bs, dcl, dtm);  // sqa_parallel_tempering.cpp:268
            bool acc = de <= 0.0;
            if (!acc) {
                if (cur >= a.cap) { ok = false; break; }
                acc = canonical_from_u64(str[cur++]) < exp(-de);
            }
            __syncthreads();
            if (acc && lane == 0)
                for (int t = 0; t < P; ++t)
                    if (incl[t]) sl(t)[i] = static_cast<int8_t>(-sl(t)[i]);
            __syncthreads();
        }
    } else {
        for (int k = 0; k < n; ++k) {
            const int i = ord[k];
            double dcl = 0.0;
            for (int t = 0; t < P; ++t) dcl += delta_fp64(a.pv, sl(t), i);
            const double de = bs * dcl;
            bool acc = de <= 0.0;
            if (!acc) {
                if (cur >= a.cap) { ok = false; break; }
                acc = canonical_from_u64(str[cur++]) < exp(-de);
            }
            __syncthreads();
            if (acc && lane == 0)
                for (int t = 0; t < P; ++t) sl(t)[i] = static_cast<int8_t>(-sl(t)[i]);
            __syncthreads();
        }
    }
    if (lane == 0) {
        a.used[r] = cur;
        if (!ok) a.status[r] = 1;
    }
}

// Integer imaginary-time bond sum B = sum_t sum_i s_t,i s_t+1,i of one state per block.
__global__ void bond_kernel(const int8_t *spins, int P, int n, int64_t *out) {
    const int8_t *S = spins + static_cast<int64_t>(blockIdx.x) * P * n;
    int64_t acc = 0;
    for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
        const int t = e / n, i = e - t * n;
        acc += static_cast<int64_t>(S[e]) * S[static_cast<int64_t>((t + 1) % P) * n + i];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ int64_t part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = 0;
        for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) s += part[w];
        out[blockIdx.x] = s;
    }
}

__global__ void fields_kernel(ProblemView pv, const int8_t *spins, double *fields, int64_t states) {
    const int64_t k = blockIdx.x;
    if (k >= states) return;
    fields_fp64(pv, spins + k * pv.n, fields + k * pv.n, threadIdx.x, blockDim.x);
}

}  // namespace

// Host: fields for `states` states in FP64.
void fields_fp64_launch(const qa_problem *p, const int8_t *d_spins, double *d_fields, int64_t states,
                        hipStream_t s) {
    hipLaunchKernelGGL(fields_kernel, dim3(static_cast<unsigned>(states)), dim3(64), 0, s, view_of(p),
                       d_spins, d_fields, states);
    QA_HIP(hipGetLastError());
}

void sa_validation(qa_ctx *ctx, const qa_problem *p, const qa_sa_params *prm, qa_sa_result *res) {
    QA_REQUIRE(prm->init_spins && prm->uniforms, "validation mode needs init_spins and uniforms");
    QA_REQUIRE(prm->uniforms_per_read >= 0, "uniforms_per_read must be >= 0");
    const int64_t n = p->n, R = prm->reads;
    for (int64_t k = 0; k < R * n; ++k)
        QA_REQUIRE(prm->init_spins[k] == 1 || prm->init_spins[k] == -1, "State spins must be -1 or +1.");
    hipStream_t s = ctx->stream;
    DevBuf<int8_t> spins(R * n), best(R * n);
    DevBuf<double> betas(prm->steps), be(R), tr(R * prm->steps), uni(std::max<int64_t>(R * prm->uniforms_per_read, 1));
    DevBuf<int64_t> used(R);
    DevBuf<int> status(R);
    DevBuf<double> rbeta(prm->read_betas ? R : 0), ienergy(prm->init_energies ? R : 0),
        fenergy(res->final_energies ? R : 0), mag(res->magnetization_trace ? R * prm->steps : 0),
        btr(res->best_energy_trace ? R * (prm->steps + 1) : 0);
    rbeta.upload(prm->read_betas, prm->read_betas ? R : 0, s);
    ienergy.upload(prm->init_energies, prm->init_energies ? R : 0, s);
    spins.upload(prm->init_spins, R * n, s);
    betas.upload(prm->betas, prm->steps, s);
    uni.upload(prm->uniforms, R * prm->uniforms_per_read, s);
    QA_HIP(hipMemsetAsync(status.ptr, 0, R * sizeof(int), s));
    QA_HIP(hipEventRecord(ctx->ev0, s));
    const int block = 64;
    hipLaunchKernelGGL(sa_validation_kernel, dim3(static_cast<unsigned>((R + block - 1) / block)), dim3(block), 0, s,
                       view_of(p), static_cast<int>(R), static_cast<int>(prm->steps), betas.ptr,
                       static_cast<int>(prm->sweeps_per_beta), spins.ptr, best.ptr, be.ptr, tr.ptr, uni.ptr,
                       prm->uniforms_per_read, used.ptr, status.ptr, rbeta.ptr, ienergy.ptr, fenergy.ptr, mag.ptr,
                       btr.ptr);
    QA_HIP(hipGetLastError());
    QA_HIP(hipEventRecord(ctx->ev1, s));
    std::vector<int> st(R);
    std::vector<int64_t> us(R);
    status.download(st.data(), R, s);
    used.download(us.data(), R, s);
    best.download(res->best_states, R * n, s);
    be.download(res->best_energies, R, s);
    tr.download(res->energy_trace, R * prm->steps, s);
    if (res->final_states) spins.download(res->final_states, R * n, s);
    if (res->final_energies) fenergy.download(res->final_energies, R, s);
    if (res->magnetization_trace) mag.download(res->magnetization_trace, R * prm->steps, s);
    if (res->best_energy_trace) btr.download(res->best_energy_trace, R * (prm->steps + 1), s);
    QA_HIP(hipStreamSynchronize(s));
    for (int64_t r = 0; r < R; ++r)
        if (st[r]) throw Error{QA_EINVAL, "validation uniform stream exhausted (uniforms_per_read too small)"};
    if (res->uniforms_used) std::memcpy(res->uniforms_used, us.data(), R * sizeof(int64_t));
    if (res->kernel_ms) {
        float ms = 0;
        QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *res->kernel_ms = ms;
    }
}

// SQA validation: per sweep phase the host shuffles each replica's persistent visit order with
// std::shuffle on that replica's engine (sqa_annealer.cpp:181,216,295), hands the device a
// chunk of raw engine outputs drawn from a copy of the engine, and afterwards advances the
// engine by exactly the number of outputs the device consumed.
// ---- SQA validation session (SQAAnnealer's sweeps, replayed phase by phase) ----------------
struct SqaValSession::Impl {
    qa_ctx *ctx;
    const qa_problem *p;
    int64_t R, P, n, cap_max;
    DevBuf<int8_t> spins, saved_spins;
    DevBuf<double> fields, saved_fields, energies;
    DevBuf<int32_t> order;
    DevBuf<int64_t> used, bonds;
    DevBuf<int> status;
    DevBuf<uint64_t> stream;
    std::vector<std::mt19937_64> eng;
    std::vector<std::vector<std::size_t>> ord;
    std::vector<uint64_t> hstream;
    std::vector<int32_t> hord;
    std::vector<int64_t> hused;
    std::vector<int> hst;
    std::vector<double> hen;
};

SqaValSession::SqaValSession(qa_ctx *ctx, const qa_problem *p, int64_t R, int64_t P, const int8_t *init_spins,
                             const uint64_t *seeds)
    : d_(std::make_shared<Impl>()) {
    QA_REQUIRE(init_spins && seeds, "validation mode needs init_spins and replica_seeds");
    QA_REQUIRE(P <= 4096, "validation mode supports at most 4096 Trotter slices");
    const int64_t n = p->n;
    for (int64_t k = 0; k < R * P * n; ++k)
        QA_REQUIRE(init_spins[k] == 1 || init_spins[k] == -1, "State spins must be -1 or +1.");
    Impl &d = *d_;
    d.ctx = ctx;
    d.p = p;
    d.R = R;
    d.P = P;
    d.n = n;
    d.cap_max = std::max<int64_t>({P * n, n * (2 * P + 2) + 64, n});
    hipStream_t s = ctx->stream;
    d.spins.alloc(R * P * n);
    d.fields.alloc(R * P * n);
    d.energies.alloc(R * P);
    d.order.alloc(R * n);
    d.used.alloc(R);
    d.bonds.alloc(R);
    d.status.alloc(R);
    d.stream.alloc(R * d.cap_max);
    d.spins.upload(init_spins, R * P * n, s);
    fields_fp64_launch(p, d.spins.ptr, d.fields.ptr, R * P, s);
    QA_HIP(hipMemsetAsync(d.status.ptr, 0, R * sizeof(int), s));
    d.eng.resize(R);
    for (int64_t r = 0; r < R; ++r) d.eng[r].seed(seeds[r]);
    d.ord.assign(R, std::vector<std::size_t>(n));
    for (auto &o : d.ord) std::iota(o.begin(), o.end(), std::size_t{0});
    d.hstream.resize(R * d.cap_max);
    d.hord.resize(R * n);
    d.hused.resize(R);
    d.hst.resize(R);
    d.hen.resize(R * P);
    best.assign(n, 1);
    best_e = std::numeric_limits<double>::infinity();
}

void SqaValSession::phase(int ph, double bs, double jp, double join) {
    Impl &d = *d_;
    const int64_t R = d.R, P = d.P, n = d.n;
    hipStream_t s = d.ctx->stream;
    const int64_t cap = ph == PH_SLICE ? P * n : (ph == PH_CLUSTER ? n * (2 * P + 2) + 64 : n);
    for (int64_t r = 0; r < R; ++r) {
        std::shuffle(d.ord[r].begin(), d.ord[r].end(), d.eng[r]);
        for (int64_t k = 0; k < n; ++k) d.hord[r * n + k] = static_cast<int32_t>(d.ord[r][k]);
        std::mt19937_64 copy = d.eng[r];
        for (int64_t k = 0; k < cap; ++k) d.hstream[r * cap + k] = copy();
    }
    d.order.upload(d.hord.data(), R * n, s);
    d.stream.upload(d.hstream.data(), R * cap, s);
    SqaPhaseArgs a{view_of(d.p), static_cast<int>(P), ph, bs, jp, join, d.spins.ptr, d.fields.ptr, d.order.ptr,
                   d.stream.ptr, cap, d.used.ptr, d.status.ptr};
    const size_t lds = ((P + 15) / 16) * 16 + P * sizeof(int);
    QA_HIP(hipEventRecord(d.ctx->ev0, s));
    hipLaunchKernelGGL(sqa_phase_kernel, dim3(static_cast<unsigned>(R)), dim3(64), lds, s, a);
    QA_HIP(hipGetLastError());
    QA_HIP(hipEventRecord(d.ctx->ev1, s));
    d.used.download(d.hused.data(), R, s);
    d.status.download(d.hst.data(), R, s);
    QA_HIP(hipStreamSynchronize(s));
    float t = 0;
    QA_HIP(hipEventElapsedTime(&t, d.ctx->ev0, d.ctx->ev1));
    ms += t;
    for (int64_t r = 0; r < R; ++r) {
        if (d.hst[r]) throw Error{QA_EDEVICE, "validation stream exhausted inside a sweep phase"};
        d.eng[r].discard(static_cast<unsigned long long>(d.hused[r]));
    }
}

void SqaValSession::bonds(std::vector<int64_t> &b) {
    Impl &d = *d_;
    hipStream_t s = d.ctx->stream;
    hipLaunchKernelGGL(bond_kernel, dim3(static_cast<unsigned>(d.R)), dim3(256), 0, s, d.spins.ptr,
                       static_cast<int>(d.P), static_cast<int>(d.n), d.bonds.ptr);
    QA_HIP(hipGetLastError());
    b.resize(d.R);
    d.bonds.download(b.data(), d.R, s);
    QA_HIP(hipStreamSynchronize(s));
}

// Step tail (sqa_annealer.cpp:455-473): FP64 slice energies on the device, then the reference's
// ordered scan for the mean and the first strict minimum.
double SqaValSession::tail() {
    Impl &d = *d_;
    hipStream_t s = d.ctx->stream;
    const int64_t RP = d.R * d.P;
    device_energies(d.p, d.spins.ptr, RP, d.energies.ptr, s);
    d.energies.download(d.hen.data(), RP, s);
    QA_HIP(hipStreamSynchronize(s));
    double avg = 0.0;
    int64_t arg = -1;
    for (int64_t k = 0; k < RP; ++k) {
        avg += d.hen[k];
        if (d.hen[k] < best_e) {
            best_e = d.hen[k];
            arg = k;
        }
    }
    if (arg >= 0) {
        QA_HIP(hipMemcpyAsync(best.data(), d.spins.ptr + arg * d.n, d.n, hipMemcpyDeviceToHost, s));
        QA_HIP(hipStreamSynchronize(s));
    }
    return avg / static_cast<double>(RP);
}

void SqaValSession::spins(int8_t *out) {
    Impl &d = *d_;
    QA_HIP(hipMemcpyAsync(out, d.spins.ptr, d.R * d.P * d.n, hipMemcpyDeviceToHost, d.ctx->stream));
    QA_HIP(hipStreamSynchronize(d.ctx->stream));
}

void SqaValSession::save() {
    Impl &d = *d_;
    const int64_t cnt = d.R * d.P * d.n;
    if (!d.saved_spins.ptr) {
        d.saved_spins.alloc(cnt);
        d.saved_fields.alloc(cnt);
    }
    QA_HIP(hipMemcpyAsync(d.saved_spins.ptr, d.spins.ptr, cnt, hipMemcpyDeviceToDevice, d.ctx->stream));
    QA_HIP(hipMemcpyAsync(d.saved_fields.ptr, d.fields.ptr, cnt * sizeof(double), hipMemcpyDeviceToDevice,
                          d.ctx->stream));
}

void SqaValSession::restore() {
    Impl &d = *d_;
    QA_REQUIRE(d.saved_spins.ptr, "no saved state");
    const int64_t cnt = d.R * d.P * d.n;
    QA_HIP(hipMemcpyAsync(d.spins.ptr, d.saved_spins.ptr, cnt, hipMemcpyDeviceToDevice, d.ctx->stream));
    QA_HIP(hipMemcpyAsync(d.fields.ptr, d.saved_fields.ptr, cnt * sizeof(double), hipMemcpyDeviceToDevice,
                          d.ctx->stream));
}

void sqa_validation(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res) {
    const int64_t P = prm->trotter_slices;
    SqaValSession v(ctx, p, prm->replicas, P, prm->init_spins, prm->replica_seeds);
    for (int64_t step = 0; step < prm->steps; ++step) {
        const double beta = prm->betas[step], gamma = prm->gammas[step];
        const double jp = trotter_coupling(beta, gamma, P);
        const double bs = beta / static_cast<double>(P);
        const double join = 1.0 - std::exp(-2.0 * jp);
        for (int64_t k = 0; k < prm->sweeps_per_beta; ++k) v.phase(PH_SLICE, bs, jp, join);
        for (int64_t k = 0; k < prm->cluster_sweeps; ++k) v.phase(PH_CLUSTER, bs, jp, join);
        for (int64_t k = 0; k < prm->worldline_sweeps; ++k) v.phase(PH_WORLDLINE, bs, jp, join);
        res->energy_trace[step] = v.tail();
        if (prm->on_step) {
            std::vector<int8_t> h(prm->replicas * P * p->n);
            v.spins(h.data());
            prm->on_step(prm->callback_user, step, res->energy_trace[step], h.data());
        }
    }
    std::memcpy(res->best_state, v.best.data(), p->n);
    *res->best_energy = v.best_e;
    if (res->kernel_ms) *res->kernel_ms = v.ms;
    if (res->accept_rate) *res->accept_rate = -1.0;
}

// ---- SQA tracing path (sqa_annealer.cpp:317-453) -------------------------------------------
// With a sweep observer the reference runs the replicas one after another from the annealer's own
// engine: slices in order, sites 0..n-1 without a shuffle, the classical dE from delta_energy (no
// field cache), then cluster and worldline sweeps, recording after every sweep. That is the SQA-PT
// rung sweep (pt_phase_kernel) on one replica with the identity visit order, fed from one engine.
void sqa_trace(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res) {
    QA_REQUIRE(prm->init_spins && prm->trace_engine, "the tracing path needs init_spins and trace_engine");
    QA_REQUIRE(res->trace_engine_out && res->trace_engine_out_len > 0, "the tracing path needs trace_engine_out");
    const int64_t n = p->n, R = prm->replicas, P = prm->trotter_slices;
    QA_REQUIRE(P <= 4096, "the tracing path supports at most 4096 Trotter slices");
    for (int64_t k = 0; k < R * P * n; ++k)
        QA_REQUIRE(prm->init_spins[k] == 1 || prm->init_spins[k] == -1, "State spins must be -1 or +1.");
    std::mt19937_64 eng;
    {
        std::istringstream in(prm->trace_engine);
        in >> eng;
        QA_REQUIRE(!in.fail(), "trace_engine is not a std::mt19937_64 state");
    }
    hipStream_t s = ctx->stream;
    const int64_t cap_max = std::max<int64_t>({P * n, n * (2 * P + 2) + 64, n});
    DevBuf<int8_t> spins(R * P * n);
    DevBuf<double> energies(R * P), dbs(1), djp(1), djoin(1);
    DevBuf<int32_t> order(n), dslot(1);
    DevBuf<int64_t> used(1);
    DevBuf<int> status(1);
    DevBuf<uint64_t> stream(cap_max);
    spins.upload(prm->init_spins, R * P * n, s);
    std::vector<int32_t> ident(n);
    std::iota(ident.begin(), ident.end(), 0);
    order.upload(ident.data(), n, s);
    QA_HIP(hipMemsetAsync(status.ptr, 0, sizeof(int), s));
    std::vector<uint64_t> hstream(cap_max);
    std::vector<double> hen(R * P), rep_e(R);
    std::vector<int8_t> host_spins(R * P * n), best(n, 1);
    double best_e = std::numeric_limits<double>::infinity();
    float total_ms = 0.0f;
    const size_t lds = ((P + 15) / 16) * 16 + P * sizeof(int);

    auto run_phase = [&](int ph, int64_t r) {
        const int64_t cap = ph == PH_SLICE ? P * n : (ph == PH_CLUSTER ? n * (2 * P + 2) + 64 : n);
        std::mt19937_64 copy = eng;
        for (int64_t k = 0; k < cap; ++k) hstream[k] = copy();
        stream.upload(hstream.data(), cap, s);
        const int32_t slot = static_cast<int32_t>(r);
        dslot.upload(&slot, 1, s);
        PtPhaseArgs a{view_of(p), static_cast<int>(P), ph, dbs.ptr, djp.ptr, djoin.ptr, dslot.ptr, spins.ptr,
                      order.ptr, stream.ptr, cap, used.ptr, status.ptr};
        QA_HIP(hipEventRecord(ctx->ev0, s));
        hipLaunchKernelGGL(pt_phase_kernel, dim3(1), dim3(64), lds, s, a);
        QA_HIP(hipGetLastError());
        QA_HIP(hipEventRecord(ctx->ev1, s));
        int64_t u = 0;
        int st = 0;
        used.download(&u, 1, s);
        status.download(&st, 1, s);
        QA_HIP(hipStreamSynchronize(s));
        float ms = 0;
        QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        total_ms += ms;
        if (st) throw Error{QA_EDEVICE, "tracing stream exhausted inside a sweep"};
        eng.discard(static_cast<unsigned long long>(u));
    };
    auto energies_now = [&]() {
        device_energies(p, spins.ptr, R * P, energies.ptr, s);
        energies.download(hen.data(), R * P, s);
        spins.download(host_spins.data(), R * P * n, s);
        QA_HIP(hipStreamSynchronize(s));
    };
    // compute_avg_energy (sqa_annealer.cpp:112-125): per-replica slice sums in slice order
    auto avg_energy = [&]() {
        double total = 0.0;
        for (int64_t r = 0; r < R; ++r) {
            double sum = 0.0;
            for (int64_t t = 0; t < P; ++t) sum += hen[r * P + t];
            rep_e[r] = sum / static_cast<double>(P);
            total += sum;
        }
        return total / static_cast<double>(R * P);
    };
    for (int64_t step = 0; step < prm->steps; ++step) {
        const double beta = prm->betas[step], gamma = prm->gammas[step];
        const double jp = trotter_coupling(beta, gamma, P);
        const double bs = beta / static_cast<double>(P);
        const double join = 1.0 - std::exp(-2.0 * jp);
        dbs.upload(&bs, 1, s);
        djp.upload(&jp, 1, s);
        djoin.upload(&join, 1, s);
        for (int64_t r = 0; r < R; ++r) {
            const struct { int ph, code; int64_t count; } phases[3] = {
                {PH_SLICE, 0, prm->sweeps_per_beta}, {PH_CLUSTER, 2, prm->cluster_sweeps},
                {PH_WORLDLINE, 1, prm->worldline_sweeps}};
            for (const auto &ph : phases)
                for (int64_t k = 0; k < ph.count; ++k) {
                    run_phase(ph.ph, r);
                    energies_now();
                    const double avg = avg_energy();
                    prm->on_sweep(prm->callback_user, step, r, k, ph.code, avg, rep_e.data(), host_spins.data());
                }
        }
        energies_now();  // step tail (:455-473)
        double avg = 0.0;
        int64_t arg = -1;
        for (int64_t k = 0; k < R * P; ++k) {
            avg += hen[k];
            if (hen[k] < best_e) {
                best_e = hen[k];
                arg = k;
            }
        }
        if (arg >= 0) std::memcpy(best.data(), host_spins.data() + arg * n, n);
        res->energy_trace[step] = avg / static_cast<double>(R * P);
        if (prm->on_step) prm->on_step(prm->callback_user, step, res->energy_trace[step], host_spins.data());
    }
    std::memcpy(res->best_state, best.data(), n);
    *res->best_energy = best_e;
    if (res->kernel_ms) *res->kernel_ms = total_ms;
    if (res->accept_rate) *res->accept_rate = -1.0;
    std::ostringstream o;
    o << eng;
    const std::string text = o.str();
    QA_REQUIRE(static_cast<int64_t>(text.size()) < res->trace_engine_out_len, "trace_engine_out is too small");
    std::memcpy(res->trace_engine_out, text.c_str(), text.size() + 1);
}

// ---- SQA-PT validation session (per-rung sweeps; states move between rungs on swaps) -------
struct PtValSession::Impl {
    qa_ctx *ctx;
    const qa_problem *p;
    int64_t R, P, n, cap_max;
    DevBuf<int8_t> spins;
    DevBuf<double> energies, dbs, djp, djoin;
    DevBuf<int32_t> order, dslot;
    DevBuf<int64_t> used, bonds;
    DevBuf<int> status;
    DevBuf<uint64_t> stream;
    std::vector<int> slot;
    std::vector<std::mt19937_64> eng;
    std::vector<std::vector<std::size_t>> ord;
    std::vector<uint64_t> hstream;
    std::vector<int32_t> hord;
    std::vector<int64_t> hused, hbond;
    std::vector<int> hst;
};

PtValSession::PtValSession(qa_ctx *ctx, const qa_problem *p, int64_t R, int64_t P, const int8_t *src,
                           bool device_src, const uint64_t *seeds)
    : d_(std::make_shared<Impl>()) {
    QA_REQUIRE(src && seeds, "validation mode needs init_spins and rung seeds");
    QA_REQUIRE(P <= 4096, "validation mode supports at most 4096 Trotter slices");
    const int64_t n = p->n;
    if (!device_src)
        for (int64_t k = 0; k < R * P * n; ++k)
            QA_REQUIRE(src[k] == 1 || src[k] == -1, "State spins must be -1 or +1.");
    Impl &d = *d_;
    d.ctx = ctx;
    d.p = p;
    d.R = R;
    d.P = P;
    d.n = n;
    d.cap_max = std::max<int64_t>({P * n, n * (2 * P + 2) + 64, n});
    hipStream_t s = ctx->stream;
    d.spins.alloc(R * P * n);
    d.energies.alloc(R * P);
    d.dbs.alloc(R);
    d.djp.alloc(R);
    d.djoin.alloc(R);
    d.order.alloc(R * n);
    d.dslot.alloc(R);
    d.used.alloc(R);
    d.bonds.alloc(R);
    d.status.alloc(R);
    d.stream.alloc(R * d.cap_max);
    if (device_src)
        QA_HIP(hipMemcpyAsync(d.spins.ptr, src, R * P * n, hipMemcpyDeviceToDevice, s));
    else
        d.spins.upload(src, R * P * n, s);
    QA_HIP(hipMemsetAsync(d.status.ptr, 0, R * sizeof(int), s));
    d.slot.resize(R);
    std::iota(d.slot.begin(), d.slot.end(), 0);
    d.eng.resize(R);
    for (int64_t r = 0; r < R; ++r) d.eng[r].seed(seeds[r]);
    d.ord.assign(R, std::vector<std::size_t>(n));
    for (auto &o : d.ord) std::iota(o.begin(), o.end(), std::size_t{0});
    d.hstream.resize(R * d.cap_max);
    d.hord.resize(R * n);
    d.hused.resize(R);
    d.hbond.resize(R);
    d.hst.resize(R);
    QA_HIP(hipStreamSynchronize(s));
}

void PtValSession::phase(int ph, const double *bs, const double *jp, const double *join) {
    Impl &d = *d_;
    const int64_t R = d.R, P = d.P, n = d.n;
    hipStream_t s = d.ctx->stream;
    const int64_t cap = ph == PH_SLICE ? P * n : (ph == PH_CLUSTER ? n * (2 * P + 2) + 64 : n);
    for (int64_t r = 0; r < R; ++r) {
        std::shuffle(d.ord[r].begin(), d.ord[r].end(), d.eng[r]);
        for (int64_t k = 0; k < n; ++k) d.hord[r * n + k] = static_cast<int32_t>(d.ord[r][k]);
        std::mt19937_64 copy = d.eng[r];
        for (int64_t k = 0; k < cap; ++k) d.hstream[r * cap + k] = copy();
    }
    d.order.upload(d.hord.data(), R * n, s);
    d.stream.upload(d.hstream.data(), R * cap, s);
    d.dslot.upload(d.slot.data(), R, s);
    d.dbs.upload(bs, R, s);
    d.djp.upload(jp, R, s);
    d.djoin.upload(join, R, s);
    PtPhaseArgs a{view_of(d.p), static_cast<int>(P), ph, d.dbs.ptr, d.djp.ptr, d.djoin.ptr, d.dslot.ptr,
                  d.spins.ptr, d.order.ptr, d.stream.ptr, cap, d.used.ptr, d.status.ptr};
    const size_t lds = ((P + 15) / 16) * 16 + P * sizeof(int);
    QA_HIP(hipEventRecord(d.ctx->ev0, s));
    hipLaunchKernelGGL(pt_phase_kernel, dim3(static_cast<unsigned>(R)), dim3(64), lds, s, a);
    QA_HIP(hipGetLastError());
    QA_HIP(hipEventRecord(d.ctx->ev1, s));
    d.used.download(d.hused.data(), R, s);
    d.status.download(d.hst.data(), R, s);
    QA_HIP(hipStreamSynchronize(s));
    float t = 0;
    QA_HIP(hipEventElapsedTime(&t, d.ctx->ev0, d.ctx->ev1));
    ms += t;
    for (int64_t r = 0; r < R; ++r) {
        if (d.hst[r]) throw Error{QA_EDEVICE, "validation stream exhausted inside a sweep phase"};
        d.eng[r].discard(static_cast<unsigned long long>(d.hused[r]));
    }
}

void PtValSession::bonds(std::vector<int64_t> &b) {
    Impl &d = *d_;
    hipStream_t s = d.ctx->stream;
    hipLaunchKernelGGL(bond_kernel, dim3(static_cast<unsigned>(d.R)), dim3(256), 0, s, d.spins.ptr,
                       static_cast<int>(d.P), static_cast<int>(d.n), d.bonds.ptr);
    QA_HIP(hipGetLastError());
    d.bonds.download(d.hbond.data(), d.R, s);
    QA_HIP(hipStreamSynchronize(s));
    b.resize(d.R);
    for (int64_t r = 0; r < d.R; ++r) b[r] = d.hbond[d.slot[r]];
}

void PtValSession::energies(std::vector<double> &e) {
    Impl &d = *d_;
    hipStream_t s = d.ctx->stream;
    const int64_t R = d.R, P = d.P;
    std::vector<double> h(R * P);
    device_energies(d.p, d.spins.ptr, R * P, d.energies.ptr, s);
    d.energies.download(h.data(), R * P, s);
    QA_HIP(hipStreamSynchronize(s));
    e.resize(R * P);
    for (int64_t r = 0; r < R; ++r)
        for (int64_t t = 0; t < P; ++t) e[r * P + t] = h[static_cast<int64_t>(d.slot[r]) * P + t];
}

void PtValSession::swap(int64_t r) { std::swap(d_->slot[r], d_->slot[r + 1]); }

void PtValSession::adopt_orders(const PtValSession &o) { d_->ord = o.d_->ord; }

void PtValSession::slice_state(int64_t r, int64_t t, int8_t *out) {
    Impl &d = *d_;
    QA_HIP(hipMemcpyAsync(out, d.spins.ptr + (static_cast<int64_t>(d.slot[r]) * d.P + t) * d.n, d.n,
                          hipMemcpyDeviceToHost, d.ctx->stream));
    QA_HIP(hipStreamSynchronize(d.ctx->stream));
}

const int8_t *PtValSession::device_spins() const { return d_->spins.ptr; }
const std::vector<int> &PtValSession::slots() const { return d_->slot; }

// SQA-PT validation: rung r keeps its engine (seeded from rung_seeds[r]) and its visit order
// while states move between rungs on accepted swaps (sqa_parallel_tempering.cpp:320-321); the
// host keeps the rung -> state-buffer map. Swap decisions use the caller's master-engine uniform
// stream, consumed only when the 4-action dE is positive.
#pragma clang fp contract(off)
void sqapt_validation(qa_ctx *ctx, const qa_problem *p, const qa_sqapt_params *prm, qa_sqapt_result *res) {
    QA_REQUIRE(prm->ladders == 1, "validation mode runs one ladder");
    QA_REQUIRE(prm->init_spins && prm->rung_seeds, "validation mode needs init_spins and rung_seeds");
    QA_REQUIRE(prm->swap_uniforms || prm->swap_uniforms_len == 0, "validation mode needs swap_uniforms");
    const int64_t n = p->n, R = prm->rungs, P = prm->trotter_slices, steps = prm->steps;
    QA_REQUIRE(P <= 4096, "validation mode supports at most 4096 Trotter slices");
    for (int64_t k = 0; k < R * P * n; ++k)
        QA_REQUIRE(prm->init_spins[k] == 1 || prm->init_spins[k] == -1, "State spins must be -1 or +1.");
    QA_REQUIRE(res->final_states && res->final_energies && res->best_state && res->best_energy &&
                   res->average_energy_trace && res->swap_acceptance_trace, "qa_sqapt_run: NULL result array");
    hipStream_t s = ctx->stream;
    DevBuf<int8_t> spins(R * P * n);
    DevBuf<double> energies(R * P), dbs(R), djp(R), djoin(R);
    DevBuf<int64_t> bonds(R), used(R);
    DevBuf<int32_t> order(R * n), dslot(R);
    DevBuf<int> status(R);
    spins.upload(prm->init_spins, R * P * n, s);
    QA_HIP(hipMemsetAsync(status.ptr, 0, R * sizeof(int), s));
    std::vector<double> hbs(R), hjp(R), hjoin(R);
    for (int64_t r = 0; r < R; ++r) {
        hbs[r] = prm->betas[r] / static_cast<double>(P);
        hjp[r] = trotter_coupling(prm->betas[r], prm->gammas[r], P);
        hjoin[r] = 1.0 - std::exp(-2.0 * hjp[r]);
    }
    dbs.upload(hbs.data(), R, s);
    djp.upload(hjp.data(), R, s);
    djoin.upload(hjoin.data(), R, s);
    std::vector<int32_t> slot(R);
    std::iota(slot.begin(), slot.end(), 0);
    std::vector<std::mt19937_64> eng(R);
    for (int64_t r = 0; r < R; ++r) eng[r].seed(prm->rung_seeds[r]);
    std::vector<std::vector<std::size_t>> ord(R, std::vector<std::size_t>(n));
    for (auto &o : ord) std::iota(o.begin(), o.end(), std::size_t{0});
    const int64_t cap_max = std::max<int64_t>({P * n, n * (2 * P + 2) + 64, n});
    DevBuf<uint64_t> stream(R * cap_max);
    std::vector<uint64_t> hstream(R * cap_max);
    std::vector<int32_t> hord(R * n);
    std::vector<int64_t> hused(R), hbond(R);
    std::vector<int> hst(R);
    std::vector<double> hen(R * P);
    int64_t ucur = 0;
    float total_ms = 0.0f;

    auto phase = [&](int ph) {
        const int64_t cap = ph == PH_SLICE ? P * n : (ph == PH_CLUSTER ? n * (2 * P + 2) + 64 : n);
        for (int64_t r = 0; r < R; ++r) {
            std::shuffle(ord[r].begin(), ord[r].end(), eng[r]);
            for (int64_t k = 0; k < n; ++k) hord[r * n + k] = static_cast<int32_t>(ord[r][k]);
            std::mt19937_64 copy = eng[r];
            for (int64_t k = 0; k < cap; ++k) hstream[r * cap + k] = copy();
        }
        order.upload(hord.data(), R * n, s);
        stream.upload(hstream.data(), R * cap, s);
        dslot.upload(slot.data(), R, s);
        PtPhaseArgs a{view_of(p), static_cast<int>(P), ph, dbs.ptr, djp.ptr, djoin.ptr, dslot.ptr, spins.ptr,
                      order.ptr, stream.ptr, cap, used.ptr, status.ptr};
        const size_t lds = ((P + 15) / 16) * 16 + P * sizeof(int);
        QA_HIP(hipEventRecord(ctx->ev0, s));
        hipLaunchKernelGGL(pt_phase_kernel, dim3(static_cast<unsigned>(R)), dim3(64), lds, s, a);
        QA_HIP(hipGetLastError());
        QA_HIP(hipEventRecord(ctx->ev1, s));
        used.download(hused.data(), R, s);
        status.download(hst.data(), R, s);
        QA_HIP(hipStreamSynchronize(s));
        float ms = 0;
        QA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        total_ms += ms;
        for (int64_t r = 0; r < R; ++r) {
            if (hst[r]) throw Error{QA_EDEVICE, "validation stream exhausted inside a sweep phase"};
            eng[r].discard(static_cast<unsigned long long>(hused[r]));
        }
    };
    auto measure = [&]() {
        device_energies(p, spins.ptr, R * P, energies.ptr, s);
        hipLaunchKernelGGL(bond_kernel, dim3(static_cast<unsigned>(R)), dim3(256), 0, s, spins.ptr,
                           static_cast<int>(P), static_cast<int>(n), bonds.ptr);
        QA_HIP(hipGetLastError());
        energies.download(hen.data(), R * P, s);
        bonds.download(hbond.data(), R, s);
        QA_HIP(hipStreamSynchronize(s));
    };
    // projected energy of the state in buffer b: first strict minimum over slices (:156-170)
    auto projected = [&](int32_t b, int64_t *best_t) {
        double m = std::numeric_limits<double>::infinity();
        int64_t bt = 0;
        for (int64_t t = 0; t < P; ++t)
            if (hen[b * P + t] < m) { m = hen[b * P + t]; bt = t; }
        *best_t = bt;
        return m;
    };
    // action (:128-154): (beta/P) sum_t E_t - J_perp B, contracted as the reference build does
    auto action = [&](int32_t b, double beta, double gamma) {
        double cs = 0.0;
        for (int64_t t = 0; t < P; ++t) cs += hen[b * P + t];
        const double jpv = trotter_coupling(beta, gamma, P);
        return std::fma(beta / static_cast<double>(P), cs, -(jpv * static_cast<double>(hbond[b])));
    };
    double best_e = std::numeric_limits<double>::infinity();
    std::vector<int8_t> best(n, 1);
    std::vector<double> en(R);
    auto take_best = [&]() {
        for (int64_t r = 0; r < R; ++r) {
            int64_t bt;
            en[r] = projected(slot[r], &bt);
            if (en[r] < best_e) {
                best_e = en[r];
                QA_HIP(hipMemcpyAsync(best.data(), spins.ptr + (static_cast<int64_t>(slot[r]) * P + bt) * n, n,
                                      hipMemcpyDeviceToHost, s));
                QA_HIP(hipStreamSynchronize(s));
            }
        }
    };
    measure();
    take_best();
    for (int64_t step = 0; step < steps; ++step) {
        for (int64_t k = 0; k < prm->sweeps_per_step; ++k) phase(PH_SLICE);
        for (int64_t k = 0; k < prm->cluster_sweeps; ++k) phase(PH_CLUSTER);
        for (int64_t k = 0; k < prm->worldline_sweeps; ++k) phase(PH_WORLDLINE);
        measure();
        take_best();
        double acc = 0.0, att = 0.0;
        if ((step + 1) % prm->swap_interval == 0) {
            for (int64_t r = 0; r + 1 < R; ++r) {
                const double eaa = action(slot[r], prm->betas[r], prm->gammas[r]);
                const double ebb = action(slot[r + 1], prm->betas[r + 1], prm->gammas[r + 1]);
                const double eab = action(slot[r + 1], prm->betas[r], prm->gammas[r]);
                const double eba = action(slot[r], prm->betas[r + 1], prm->gammas[r + 1]);
                const double d = (eab + eba) - (eaa + ebb);
                att += 1.0;
                bool ok = d <= 0.0;
                if (!ok) {
                    if (ucur >= prm->swap_uniforms_len) throw Error{QA_EINVAL, "swap uniform stream exhausted"};
                    ok = prm->swap_uniforms[ucur++] < std::exp(-d);
                }
                if (ok) {
                    std::swap(slot[r], slot[r + 1]);
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
    for (int64_t r = 0; r < R; ++r) {
        int64_t bt;
        res->final_energies[r] = projected(slot[r], &bt);
        QA_HIP(hipMemcpyAsync(res->final_states + r * n, spins.ptr + (static_cast<int64_t>(slot[r]) * P + bt) * n, n,
                              hipMemcpyDeviceToHost, s));
    }
    QA_HIP(hipStreamSynchronize(s));
    std::memcpy(res->best_state, best.data(), n);
    *res->best_energy = best_e;
    if (res->swap_uniforms_used) *res->swap_uniforms_used = ucur;
    if (res->kernel_ms) *res->kernel_ms = total_ms;
}
#pragma clang fp contract(on)

}  // namespace qa
