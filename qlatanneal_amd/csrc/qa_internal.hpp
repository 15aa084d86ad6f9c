// Internal declarations of libqanneal_hip.so (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <tuple>

#include <cstdint>
#include <string>
#include <vector>

#include "qanneal_hip.h"

namespace qa {

// ---- error plumbing ------------------------------------------------------------------------
void set_error(const std::string &msg);
// Raises a kernel's dynamic-LDS limit (hipFuncAttributeMaxDynamicSharedMemorySize) to at least
// `bytes` on the current device. The attribute belongs to the kernel on that device, for the whole
// process: here it only ever grows, under one lock, so runs on concurrent host threads never lower
// it below a size another thread is about to launch with, and a launch whose size is already
// covered on its device makes no runtime call.
void ensure_dynamic_lds(const void *kern, size_t bytes);

struct Error {
    qa_status status;
    std::string msg;
};

#define QA_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::qa::Error{e_ == hipErrorOutOfMemory ? QA_ENOMEM : QA_EDEVICE,          \
                              std::string(#call) + ": " + hipGetErrorString(e_)};         \
    } while (0)

#define QA_REQUIRE(cond, msg)                                                              \
    do {                                                                                   \
        if (!(cond)) throw ::qa::Error{QA_EINVAL, (msg)};                                  \
    } while (0)

// ---- device buffers ------------------------------------------------------------------------
// hipFree synchronizes the whole device, so a run that freed its scratch with it would wait for
// the kernels of every other stream. While a run is active on a thread (RunScope), device buffers
// come from and go back to the cache of the run's context instead. A context's buffers are only
// handed out again to work of that context, i.e. on its one stream, so stream order makes a
// reused buffer safe without any host synchronization. Outside a run (problem uploads and other
// long-lived buffers) allocation is plain hipMalloc / hipFree.
struct BufCache;
extern thread_local BufCache *t_run_cache;
void *cache_alloc(BufCache *c, size_t bytes, size_t *cls);  // cls: 0 = not cached (plain hipMalloc)
void cache_release(BufCache *c, void *ptr, size_t cls);
BufCache *cache_create();
void cache_destroy(BufCache *c);  // frees the cached buffers; live ones are freed on release

template <class T>
struct DevBuf {
    T *ptr = nullptr;
    size_t count = 0;
    BufCache *cache = nullptr;
    size_t cls = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), count(o.count), cache(o.cache), cls(o.cls) {
        o.ptr = nullptr;
        o.count = 0;
        o.cache = nullptr;
        o.cls = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            count = o.count;
            cache = o.cache;
            cls = o.cls;
            o.ptr = nullptr;
            o.count = 0;
            o.cache = nullptr;
            o.cls = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        count = n;
        if (!n) return;
        if (t_run_cache) {
            ptr = static_cast<T *>(cache_alloc(t_run_cache, n * sizeof(T), &cls));
            cache = cls ? t_run_cache : nullptr;
        } else {
            QA_HIP(hipMalloc(&ptr, n * sizeof(T)));
        }
    }
    void release() {
        if (ptr) {
            if (cache) cache_release(cache, ptr, cls);
            else (void)hipFree(ptr);
        }
        ptr = nullptr;
        count = 0;
        cache = nullptr;
        cls = 0;
    }
    void upload(const T *src, size_t n, hipStream_t s) {
        if (n) QA_HIP(hipMemcpyAsync(ptr, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *dst, size_t n, hipStream_t s) const {
        if (n) QA_HIP(hipMemcpyAsync(dst, ptr, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

}  // namespace qa

struct qa_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    qa::BufCache *cache = nullptr;  // run-scoped device buffers of this context's stream
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int num_cus = 256;
    size_t lds_per_block = 65536;
    // batch autotuner: replicas per group chosen for (model kind, n rounded up to 64, P)
    std::map<std::tuple<int, int64_t, int64_t>, int64_t> tuned_rg;
};

namespace qa {
// Marks the calling thread as running work of `ctx` (see DevBuf); scopes nest.
struct RunScope {
    BufCache *prev;
    explicit RunScope(const qa_ctx *ctx) : prev(t_run_cache) { t_run_cache = ctx ? ctx->cache : nullptr; }
    ~RunScope() { t_run_cache = prev; }
    RunScope(const RunScope &) = delete;
    RunScope &operator=(const RunScope &) = delete;
};
}  // namespace qa

enum { QA_KIND_DENSE = 0, QA_KIND_SPARSE = 1, QA_KIND_HUBO = 2 };

// Device-resident problem. FP64 copies serve validation runs and exact energies; FP32 copies
// (diagonal of J cleared: the dynamics never read it) serve the production sweep.
struct qa_problem {
    qa_ctx *ctx = nullptr;
    int kind = QA_KIND_DENSE;
    int64_t n = 0;
    double c = 0.0;
    bool symmetric = true;
    qa::DevBuf<double> h64;      // [n]
    qa::DevBuf<float> h32;       // [n]
    // dense
    qa::DevBuf<double> J64;      // [n*n] as given
    qa::DevBuf<float> J32;       // [n*n] zero diagonal
    qa::DevBuf<float> J32p;      // [n][n rounded up to 4], zero padded; only when n % 4 != 0
    qa::DevBuf<uint16_t> Jbf_hi, Jbf_lo;  // bf16 split of J32 for the matrix-core field refresh
    int ldk_bf = 0;                        //   [n rounded up to 64][ldk_bf = n rounded up to 16]
    // sparse: undirected edges (energy order) + adjacency CSR in reference insertion order
    int64_t nedges = 0;
    qa::DevBuf<int32_t> ei, ej;
    qa::DevBuf<double> ev64;
    qa::DevBuf<int32_t> adj_ptr, adj_idx;  // [n+1], [2*nedges]
    bool adj_unique = true;                // no spin appears twice in one adjacency row
    qa::DevBuf<double> adj_val64;
    qa::DevBuf<float> adj_val32;
    // hubo: consolidated terms + incidence CSR (term ids per variable, ascending)
    int64_t nterms = 0;
    qa::DevBuf<int32_t> t_off, t_vars;
    qa::DevBuf<double> t_coef64;
    qa::DevBuf<float> t_coef32;
    qa::DevBuf<int32_t> inc_ptr, inc_idx;
    int max_degree = 0;          // sparse: max adjacency length; hubo: max incident terms
    // instance statistics for the chi_B surrogate's features (accumulated at upload)
    double st_mean_degree = 0.0;  // couplings per spin (hubo: sum over incident terms of order - 1)
    double st_field_scale = 0.0;  // mean over spins of sqrt(sum of squared incident couplings)
    double st_field_het = 0.0;    // std / mean of that per-spin scale
    double st_kurtosis = 1.0;     // E[J^4] / E[J^2]^2 over the nonzero couplings
    double st_neg_frac = 0.0;     // fraction of negative couplings
    double st_h_ratio = 0.0;      // mean |h| / field scale
    double st_order = 2.0;        // mean interaction order
};

// chi_B surrogate network (an ensemble of `members` MLPs sharing one launch per layer)
struct qa_mlp {
    qa_ctx *ctx = nullptr;
    int n_in = 0, width = 0, depth = 0, n_out = 0, members = 0, maxdim = 0;
    int64_t nparams = 0;                 // per member
    std::vector<int64_t> w_off, b_off;   // per layer, within a member
    qa::DevBuf<float> params, grads, m1, m2;  // [members][nparams] (FP32 master copy, Adam state)
    qa::DevBuf<uint16_t> wbf;            // [members][nparams] bf16 copy the products read
    qa::DevBuf<double> loss;             // [members]
    std::vector<double> x_mean, x_scale; // input normalization
    std::vector<double> x_lo, x_hi;      // training range of the normalized inputs
    bool trained = false;
    int64_t adam_t = 0;
    int64_t cap = 0;                     // batch capacity of the work buffers
    qa::DevBuf<float> acts;              // [members][depth+1][cap][maxdim] layer outputs
    qa::DevBuf<float> dA, dB;            // [members][cap][maxdim] gradients
    qa::DevBuf<float> xin, tgt;          // [cap][n_in], [cap][n_out]
};

// Views passed to kernels (by value).
struct ProblemView {
    int kind;
    int n;
    double c;
    const double *h64;
    const float *h32;
    const double *J64;
    const float *J32;
    const float *J32p;  // J32 with rows padded to a multiple of 4 floats (J32 itself when n % 4 == 0)
    int nedges;
    const int32_t *ei, *ej;
    const double *ev64;
    const int32_t *adj_ptr, *adj_idx;
    const double *adj_val64;
    const float *adj_val32;
    int nterms;
    const int32_t *t_off, *t_vars;
    const double *t_coef64;
    const float *t_coef32;
    const int32_t *inc_ptr, *inc_idx;
};

inline ProblemView view_of(const qa_problem *p) {
    ProblemView v{};
    v.kind = p->kind;
    v.n = static_cast<int>(p->n);
    v.c = p->c;
    v.h64 = p->h64.ptr;
    v.h32 = p->h32.ptr;
    v.J64 = p->J64.ptr;
    v.J32 = p->J32.ptr;
    v.J32p = (p->n % 4 == 0) ? p->J32.ptr : p->J32p.ptr;
    v.nedges = static_cast<int>(p->nedges);
    v.ei = p->ei.ptr;
    v.ej = p->ej.ptr;
    v.ev64 = p->ev64.ptr;
    v.adj_ptr = p->adj_ptr.ptr;
    v.adj_idx = p->adj_idx.ptr;
    v.adj_val64 = p->adj_val64.ptr;
    v.adj_val32 = p->adj_val32.ptr;
    v.nterms = static_cast<int>(p->nterms);
    v.t_off = p->t_off.ptr;
    v.t_vars = p->t_vars.ptr;
    v.t_coef64 = p->t_coef64.ptr;
    v.t_coef32 = p->t_coef32.ptr;
    v.inc_ptr = p->inc_ptr.ptr;
    v.inc_idx = p->inc_idx.ptr;
    return v;
}

namespace qa {

// Trotter coupling J_perp = 0.5 ln(1 / tanh(max(beta*Gamma/P, 1e-12))) — the formula of
// sqa_annealer.cpp:66-70, evaluated on the host in FP64 for every step.
double trotter_coupling(double beta, double gamma, int64_t slices);

// Energies on the device (FP64, reference order) for `count` states already resident.
void device_energies(const qa_problem *p, const int8_t *d_spins, int64_t count, double *d_out,
                     hipStream_t s);

// Matrix-core field refresh (qa_fields.hip): bf16 split of the FP32 couplings (dense, symmetric);
// FP32 fields of `count` resident states (row stride ldf; models other than symmetric dense go
// through an FP64 scratch of count*n doubles); and the max |FP32 - FP64| field error of resident
// states against an FP64 recompute in the reference order, folded into *d_max as the bits of a
// non-negative double (read with drift_value). All launches are asynchronous on stream s.
void build_bf16_split(qa_problem *p, const float *J32, hipStream_t s);
void fields_fp32_launch(const qa_problem *p, const int8_t *d_spins, int64_t count, float *d_fields, int64_t ldf,
                        double *d_scratch, hipStream_t s);
void field_drift_launch(const qa_problem *p, const int8_t *d_spins, const float *d_fields, int64_t ldf, int64_t count,
                        double *d_scratch, unsigned long long *d_max, hipStream_t s);
double drift_value(unsigned long long bits);
void fields_fp64_launch(const qa_problem *p, const int8_t *d_spins, double *d_fields, int64_t states,
                        hipStream_t s);

// Engine entry points (implemented per engine file).
void sqa_optimal(qa_ctx *ctx, const qa_problem *p, const qa_sqa_opt_params *prm, qa_sqa_opt_result *res);
void sqapt_optimal(qa_ctx *ctx, const qa_problem *p, const qa_sqapt_opt_params *prm, qa_sqapt_opt_result *res);
void sa_validation(qa_ctx *ctx, const qa_problem *p, const qa_sa_params *prm, qa_sa_result *res);
void sa_fp32(qa_ctx *ctx, const qa_problem *p, const qa_sa_params *prm, qa_sa_result *res);
void sqa_validation(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res);
void sqa_trace(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res);
void sqa_fp32(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res);
void sqa_fp32_noalloc(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm, qa_sqa_result *res);
void sqapt_validation(qa_ctx *ctx, const qa_problem *p, const qa_sqapt_params *prm, qa_sqapt_result *res);
void sqapt_fp32(qa_ctx *ctx, const qa_problem *p, const qa_sqapt_params *prm, qa_sqapt_result *res);
void chi_validation(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, qa_chi_result *res);
void chi_fp32(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, qa_chi_result *res);
void chi_fp32_schedule(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, qa_chi_result *res);
// instance statistics from per-spin coupling sums (see qa_problem::st_*)
struct StatsAcc {
    std::vector<double> deg, r2;
    double s2 = 0.0, s4 = 0.0, cnt = 0.0, neg = 0.0, order_sum = 0.0;
    explicit StatsAcc(int64_t n) : deg(static_cast<size_t>(n), 0.0), r2(static_cast<size_t>(n), 0.0) {}
    void coupling(double v, int order) {
        s2 += v * v;
        s4 += v * v * v * v;
        cnt += 1.0;
        if (v < 0.0) neg += 1.0;
        order_sum += order;
    }
    void finish(qa_problem *p, const double *h) const;
};
void chi_features(const qa_problem *p, double beta, int64_t P, double A0, double gamma_start, double gamma_end,
                  double *f);
void chi_pilot_profile(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, double *scan_s, double *scan_gamma,
                       double *scan_chi, float *ms);
void mlp_init(qa_mlp *m, uint64_t seed);
void mlp_sync_bf16(qa_mlp *m);
void mlp_forward(qa_mlp *m, int64_t B, const double *x, float *y);
void mlp_loss_grad(qa_mlp *m, int64_t B, const double *x, const float *t, double *loss, float *grad);
void mlp_train(qa_mlp *m, int64_t S, const double *x, const float *t, const qa_mlp_train_params *tp, double *history);
void sqa_batch(qa_ctx *ctx, int64_t count, const qa_problem *const *problems, const qa_batch_params *prm,
               qa_batch_result *res);

// ---- adaptive J_perp schedules (run_optimal) -----------------------------------------------
// B statistics of one step: sums of B and B^2 over every replica's samples (one sample after
// each slice and each cluster sweep; B = sum_t sum_i s_t,i s_t+1,i is an integer, so the sums
// are exact), the sample count, and B of replica 0 (buffer 0) at the end of the step.
struct BondStats {
    double sum = 0.0, sum2 = 0.0;
    int64_t count = 0;
    double b0 = 0.0;
};

// FP32 SQA state kept on the device across launches. run() sweeps k steps with per-step
// (beta/P, J_perp), or one step with per-replica parameters (param_of_rep[r] indexes bs/jp);
// tail 0 records the step ends (best state, mean slice energy), tail 1 leaves per-chain slice
// energies in chain_e (FP64 sums over the FP32 fields), tail 2 records nothing. The Philox
// streams continue from launch to launch.
class SqaFp32Session {
public:
    SqaFp32Session(qa_ctx *ctx, const qa_problem *p, int64_t R, int64_t P, uint64_t key);
    void copy_state_from(const SqaFp32Session &o);
    void run(int64_t k, const double *bs, const double *jp, int64_t nparam, const int *param_of_rep, int64_t spb,
             int64_t cl, int64_t wl, int tail, BondStats *stats, double *trace);
    void save();
    void restore();
    // best slice state over the tail-0 steps and its FP64 energy recomputed on the device
    void best(int8_t *state, double *energy);
    const int8_t *device_spins() const { return gS_.ptr; }
    std::vector<double> chain_e;  // tail 1: [R*P]
    std::vector<int> chain_b;     // tail 1: [R*P] sum_i s_t,i s_t+1,i per chain
    float ms = 0.0f;

private:
    qa_ctx *ctx_;
    const qa_problem *p_;
    int64_t R_, P_, rg_, ngroups_;
    bool use_lds_;
    uint64_t key_;
    uint32_t sweeps_ = 0;
    int64_t recorded_ = 0;
    bool started_ = false;
    DevBuf<int8_t> gS_, sS_, bstate_;
    DevBuf<float> gF_, sF_;
    DevBuf<double> be_, ce_;
    DevBuf<int> bstep_, bchain_, cb_;
};

// FP64 validation replay of SQAAnnealer's sweeps (sqa_annealer.cpp:176-315) with the caller's
// per-replica engines; the host draws the visit orders and uniform streams and advances the
// engines by what the device consumed.
class SqaValSession {
public:
    SqaValSession(qa_ctx *ctx, const qa_problem *p, int64_t R, int64_t P, const int8_t *init_spins,
                  const uint64_t *seeds);
    void phase(int ph, double bs, double jp, double join);
    void bonds(std::vector<int64_t> &b);  // B of every replica from the current spins
    double tail();                        // step end: best update (first strict minimum), mean energy
    void spins(int8_t *out);              // host copy of the lattice [R*P*n]
    void save();
    void restore();
    std::vector<int8_t> best;
    double best_e;
    float ms = 0.0f;
    static constexpr int kSlice = 0, kCluster = 1, kWorldline = 2;

private:
    struct Impl;
    std::shared_ptr<Impl> d_;
};

// FP64 validation replay of SQAParallelTemperingAnnealer's per-rung sweeps
// (sqa_parallel_tempering.cpp:208-296): rung r keeps its engine and visit order while states
// move between rungs on swaps (slot[r] = buffer holding rung r).
class PtValSession {
public:
    PtValSession(qa_ctx *ctx, const qa_problem *p, int64_t R, int64_t P, const int8_t *d_or_h_spins, bool device_src,
                 const uint64_t *seeds);
    void phase(int ph, const double *bs, const double *jp, const double *join);
    void bonds(std::vector<int64_t> &b);    // by rung
    void energies(std::vector<double> &e);  // [rung][slice] FP64 slice energies
    void swap(int64_t r);                   // rungs r and r + 1 exchange states
    void adopt_orders(const PtValSession &o);  // take o's per-rung visit orders (shared in the reference)
    void slice_state(int64_t r, int64_t t, int8_t *out);
    const int8_t *device_spins() const;
    const std::vector<int> &slots() const;
    float ms = 0.0f;

private:
    struct Impl;
    std::shared_ptr<Impl> d_;
};

}  // namespace qa
