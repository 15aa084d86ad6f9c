/*
 * qanneal_hip.h — C-ABI of the gfx950 annealing-sweep library (qlatanneal_amd/libqanneal_hip.so).
 *
 * This is the device boundary of the annealing-sweep path. The reference reaches its engines
 * through per-spin virtual calls (include/qanneal/backend.hpp:39-53) from inside the run loops;
 * here the boundary sits one level up, at whole runs: the C++ annealer classes
 * (the include/qanneal headers of this repository) hand a complete run to one of the qa_*_run entry
 * points. Signatures use plain pointers and sizes only. Every buffer is caller-allocated host
 * memory; the library never returns memory it allocated.
 *
 * Modes. QA_MODE_FP32 is the production sweep (FP32 local fields, Philox-4x32-10 counter RNG,
 * lockstep replica groups). QA_MODE_FP64_VALIDATION reproduces the reference CPU trajectory: FP64
 * fields and energies in the reference's summation order, and the reference's random stream
 * (std::mt19937_64 draws, consumed exactly where the reference consumes them).
 *
 * Threading: calls on one qa_ctx are serialized by the caller (one context per host thread);
 * every call blocks until its device work has completed.
 */
#ifndef QANNEAL_HIP_H
#define QANNEAL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum qa_status {
    QA_OK = 0,
    QA_EINVAL = 1,        /* invalid argument (reference: std::invalid_argument)        */
    QA_ENOMEM = 2,        /* device allocation failed                                    */
    QA_EDEVICE = 3,       /* HIP runtime / kernel failure (reference: std::runtime_error) */
    QA_EUNSUPPORTED = 4   /* valid request this build does not run                       */
} qa_status;

typedef enum qa_mode { QA_MODE_FP32 = 0, QA_MODE_FP64_VALIDATION = 1 } qa_mode;

typedef struct qa_ctx qa_ctx;
typedef struct qa_problem qa_problem;

/* Message for the last failing call on this host thread. */
const char *qa_last_error(void);

/* Replaces make_backend(kind, ham) (backend.hpp:92-110) for the device part: binds a GPU. */
qa_status qa_ctx_create(int device, qa_ctx **out);
qa_status qa_ctx_destroy(qa_ctx *ctx);

/* Device copies of the problem models. Each replaces the Hamiltonian object the reference's
 * CPUBackend wraps (backend.hpp:55-90):
 *   dense  — DenseIsing(h, J, n, c)          dense_ising.hpp:17-21; J row-major n*n
 *   sparse — SparseIsing(h, edges, n, c)     sparse_ising.hpp:21-25; edges in caller order
 *   hubo   — HigherOrderIsing after its ctor's consolidation (higher_order_ising.cpp:10-90):
 *            linear h[n], terms as CSR (term_offset[nterms+1], term_vars), term_coeff[nterms].
 * The problem is uploaded once and reused by every run. */
qa_status qa_problem_dense(qa_ctx *ctx, int64_t n, const double *h, const double *J, double c,
                           qa_problem **out);
qa_status qa_problem_sparse(qa_ctx *ctx, int64_t n, const double *h, int64_t nedges,
                            const int64_t *ei, const int64_t *ej, const double *ev, double c,
                            qa_problem **out);
qa_status qa_problem_hubo(qa_ctx *ctx, int64_t n, const double *h, int64_t nterms,
                          const int64_t *term_offset, const uint32_t *term_vars,
                          const double *term_coeff, double c, qa_problem **out);
qa_status qa_problem_destroy(qa_problem *p);

/* FP64 energies of `count` states (spins [count*n] in {-1,+1}), computed on the device in the
 * reference's summation order, so the result equals Hamiltonian::energy bit for bit
 * (dense_ising.cpp:29-44, sparse_ising.cpp:43-56, higher_order_ising.cpp:92-111). */
qa_status qa_energy(const qa_problem *p, int64_t count, const int8_t *spins, double *out);

/* FP32 local fields f_i = h_i + sum_{j != i} J_ij s_j of `count` states (spins [count*n]) into
 * fields[count*n]: DenseIsing::compute_local_fields (dense_ising.cpp:64-83) for many states at
 * once. Symmetric dense couplings run on the matrix cores (bf16 hi/lo split of the FP32 J, FP32
 * accumulation; couplings with <= 16 significant bits are exact); other models round the FP64
 * fields of the reference order (sparse_ising.cpp:73-85, higher_order_ising.cpp:138-160). */
qa_status qa_local_fields(const qa_problem *p, int64_t count, const int8_t *spins, float *fields);

/* ---- SA: replaces Annealer::run (annealer.cpp:33-75), `reads` independent chains ---------- */
typedef struct qa_sa_params {
    int64_t steps;               /* schedule length                                          */
    const double *betas;         /* [steps]                                                  */
    int64_t sweeps_per_beta;
    int64_t reads;               /* independent chains run together                          */
    int32_t mode;                /* qa_mode                                                  */
    uint64_t seed;               /* FP32: Philox key                                         */
    const int8_t *init_spins;    /* validation: [reads*n], State::random draws               */
    const double *uniforms;      /* validation: [reads*uniforms_per_read], the engine's
                                    uniform_real(0,1) stream after the init draws            */
    int64_t uniforms_per_read;
    /* Optional, both modes: chained runs (ParallelTemperingAnnealer::run, parallel_tempering.cpp:
     * 36-122, runs each replica at its own beta and continues its state and running energy). */
    const double *read_betas;    /* [reads]: read r uses read_betas[r] at every step (betas unused) */
    const int8_t *init_states;   /* FP32: [reads*n] start states instead of random ones        */
    const double *init_energies; /* [reads]: running energy at the start (default: the FP64
                                    energy of the start state)                               */
    uint64_t counter_base;       /* FP32: Philox stream offset of this run (chained runs)      */
} qa_sa_params;

typedef struct qa_sa_result {
    int8_t *best_states;         /* [reads*n]                                                */
    double *best_energies;       /* [reads]                                                  */
    double *energy_trace;        /* [reads*steps] energy at the end of each step             */
    int64_t *uniforms_used;      /* validation: [reads] (may be NULL)                        */
    double *kernel_ms;           /* optional: device time of the sweep kernel                */
    int8_t *final_states;        /* optional [reads*n]: states at the end of the run          */
    double *final_energies;      /* optional [reads]: energies at the end of the run; FP32 mode:
                                    the FP64 energy of the final state (recomputed on the
                                    device), validation: the running energy                  */
    double *magnetization_trace; /* optional [reads*steps]: mean spin at each step end
                                    (metrics.hpp:11-20; ReplicaAnnealer::run's trace)        */
    double *best_energy_trace;   /* optional [reads*(steps+1)]: best running energy at the start
                                    and at each step end (replays ReplicaAnnealer::run's
                                    global-best updates, replica_annealer.cpp:111-116)       */
} qa_sa_result;

qa_status qa_sa_run(qa_ctx *ctx, const qa_problem *p, const qa_sa_params *prm, qa_sa_result *res);

/* ---- SQA: replaces SQAAnnealer::run (sqa_annealer.cpp:87-477) ----------------------------- */
/* Observer callbacks (SQAAnnealer::run's observer calls, sqa_annealer.cpp:340-350, 474-476). They run
 * on the calling thread between launches; `spins` is a host copy of the whole [R*P*n] lattice,
 * valid during the call. phase: 0 slice, 1 worldline, 2 cluster (the reference's SQASweepPhase). */
typedef void (*qa_sqa_step_fn)(void *user, int64_t step, double avg_energy, const int8_t *spins);
typedef void (*qa_sqa_sweep_fn)(void *user, int64_t step, int64_t replica, int64_t sweep, int32_t phase,
                                double avg_energy, const double *replica_energies, const int8_t *spins);

typedef struct qa_sqa_params {
    int64_t steps;
    const double *betas;         /* [steps]                                                  */
    const double *gammas;        /* [steps]                                                  */
    int64_t trotter_slices;      /* P actually used (caller resolves continuous_time_slices) */
    int64_t replicas;
    int64_t sweeps_per_beta;
    int64_t worldline_sweeps;
    int64_t cluster_sweeps;
    int32_t mode;                /* qa_mode                                                  */
    uint64_t seed;               /* FP32: Philox key                                         */
    const int8_t *init_spins;    /* validation: [replicas*P*n], SQAState::random draws       */
    const uint64_t *replica_seeds; /* validation: [replicas], the per-replica engine seeds   */
    int64_t refresh_interval;    /* FP32: sweeps between recomputations of all FP32 fields on
                                    the matrix cores from the spins (0: never; no reference
                                    counterpart, the reference keeps FP64 fields)             */
    qa_sqa_step_fn on_step;      /* optional: after every step (the run goes launch by launch;
                                    same trajectory; not with refresh_interval)               */
    qa_sqa_sweep_fn on_sweep;    /* optional: selects the reference's sequential tracing path
                                    (sqa_annealer.cpp:317-453), FP64 on the device in either
                                    mode; needs init_spins and trace_engine                    */
    void *callback_user;
    const char *trace_engine;    /* tracing path: the annealer's std::mt19937_64 after the
                                    replica seeds, in the standard's text form (operator<<)   */
} qa_sqa_params;

typedef struct qa_sqa_result {
    int8_t *best_state;          /* [n]                                                      */
    double *best_energy;         /* FP64 energy of best_state (recomputed on the device)     */
    double *energy_trace;        /* [steps] mean slice energy at each step end               */
    double *kernel_ms;           /* optional: device time of the sweep kernel(s)             */
    double *accept_rate;         /* optional: accepted / attempted slice moves (FP32 mode)   */
    double *field_drift;         /* optional [2], FP32 with refresh_interval > 0: max |f32 -
                                    f64| over up to 64 chains before each refresh, and after
                                    it (f64: the reference-order FP64 fields); NaN if none    */
    char *trace_engine_out;      /* tracing path: the engine after the run, in text form       */
    int64_t trace_engine_out_len;/* capacity of trace_engine_out (8192 is enough)             */
} qa_sqa_result;

qa_status qa_sqa_run(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm,
                     qa_sqa_result *res);

/* FP32 SQA run that allocates nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (serving loops, harnesses that forbid hipMalloc). Same run and
 * same results as qa_sqa_run in QA_MODE_FP32 (bit for bit) for refresh_interval = 0 and no
 * callbacks; every buffer beyond the problem and the caller's result arrays lives in a static
 * per-device scratch array of 64 MiB (spins R*P*n bytes, fields R*P*n*4 bytes when they do not
 * fit LDS, per-group best states and trace partials, the schedule). A run that needs more is
 * refused with QA_EINVAL before anything is launched. Calls on one device share that array and
 * are serialized by the library; each call overwrites the previous call's contents. */
qa_status qa_sqa_run_noalloc(qa_ctx *ctx, const qa_problem *p, const qa_sqa_params *prm,
                             qa_sqa_result *res);

/* ---- Batched SQA: many problems of uneven size in one persistent launch ---------------------
 * No reference counterpart (the reference runs instances one by one, solve() in solver.py:455-789).
 * Instance k runs what qa_sqa_run runs in FP32 mode with replicas = reads[k] and seed = seeds[k],
 * all instances sharing the schedule. Work items (instance, replica group) are dealt to persistent
 * workgroups through an atomic counter, largest first; one kernel per sweep variant. With
 * autotune = 0 every instance's replica groups are those of its own qa_sqa_run call and its result
 * equals that call's bit for bit; autotune = 1 picks replicas per group by timed trials, cached per
 * (model, size class, P) in the context. use_graph = 1 captures the launch sequence in a hipGraph.
 * Status: switched off. The launch did not complete on its first MI355X run and no cause has been
 * shown, so the call validates its arguments and returns QA_EUNSUPPORTED without launching. */
typedef struct qa_batch_params {
    int64_t steps;
    const double *betas;         /* [steps], shared by all instances                         */
    const double *gammas;        /* [steps]                                                  */
    int64_t trotter_slices;
    int64_t sweeps_per_beta;
    int64_t worldline_sweeps;
    int64_t cluster_sweeps;
    const int64_t *reads;        /* [count] replicas per instance                            */
    const uint64_t *seeds;       /* [count] Philox keys                                      */
    int32_t autotune;
    int32_t use_graph;
} qa_batch_params;

typedef struct qa_batch_result {
    int8_t *best_states;         /* [sum_k n_k]: instance k at offset n_0 + ... + n_{k-1}    */
    double *best_energies;       /* [count] FP64 energies of the best states                 */
    double *kernel_ms;           /* optional: device time of the batch launch sequence       */
    int64_t *attempts;           /* optional: single-spin attempts of the whole batch        */
} qa_batch_result;

qa_status qa_sqa_batch_run(qa_ctx *ctx, int64_t count, const qa_problem *const *problems,
                           const qa_batch_params *prm, qa_batch_result *res);

/* ---- SQA parallel tempering: replaces SQAParallelTemperingAnnealer::run
 *      (sqa_parallel_tempering.cpp:81-346) ------------------------------------------------- */
typedef struct qa_sqapt_params {
    int64_t rungs;               /* ladder length                                            */
    const double *betas;         /* [rungs]                                                  */
    const double *gammas;        /* [rungs]                                                  */
    int64_t ladders;             /* independent ladders run together (FP32 mode)             */
    int64_t trotter_slices;
    int64_t sweeps_per_step;
    int64_t worldline_sweeps;
    int64_t cluster_sweeps;
    int64_t steps;
    int64_t swap_interval;
    int32_t mode;
    uint64_t seed;
    const int8_t *init_spins;    /* validation: [rungs*P*n]                                  */
    const uint64_t *rung_seeds;  /* validation: [rungs]                                      */
    const double *swap_uniforms; /* validation: master-engine uniform stream for swaps       */
    int64_t swap_uniforms_len;
    int64_t refresh_interval;    /* FP32: as in qa_sqa_params (sweeps between field refreshes) */
    /* FP32 mode, optional (NULL: the fixed ladder betas/gammas at every step): a ladder per PT
     * step, [steps*rungs] each. Rung r sweeps step k at (step_betas[k*rungs + r],
     * step_gammas[k*rungs + r]) and the exchange after step k is decided with the same pairs, so
     * every rung anneals (an annealed ladder, e.g. on the schedule qa_chi_schedule returns) while
     * replicas still move between neighbouring rungs. The reference's ladder is fixed for a run
     * (sqa_parallel_tempering.cpp:97-200); with every row equal to betas/gammas the run is
     * bit-identical to the fixed-ladder run. betas/gammas must still hold a valid ladder. */
    const double *step_betas, *step_gammas;
} qa_sqapt_params;

typedef struct qa_sqapt_result {
    int8_t *final_states;        /* [ladders*rungs*n] best slice per rung at the end         */
    double *final_energies;      /* [ladders*rungs]                                          */
    int8_t *best_state;          /* [ladders*n]                                              */
    double *best_energy;         /* [ladders]                                                */
    double *average_energy_trace;/* [ladders*steps]                                          */
    double *swap_acceptance_trace;/* [ladders*steps]                                         */
    int64_t *swap_uniforms_used; /* validation (may be NULL)                                 */
    double *kernel_ms;
    double *field_drift;         /* optional [2]: as in qa_sqa_result                         */
    double *sweep_ms;            /* optional (FP32): device time of the sweep launches alone,
                                    summed over the run (kernel_ms adds the exchange, refresh
                                    and drift kernels)                                        */
    double *accept_rate;         /* optional (FP32): flipped (chain, site) visits over all
                                    visits of all sweeps of the run                           */
} qa_sqapt_result;

qa_status qa_sqapt_run(qa_ctx *ctx, const qa_problem *p, const qa_sqapt_params *prm,
                       qa_sqapt_result *res);

/* ---- Adaptive transverse-coupling schedules -------------------------------------------------
 * The bond susceptibility chi_B = Var(B), B = sum_t sum_i s_t,i s_t+1,i, sampled after every
 * slice and cluster sweep, sets how fast J_perp advances: a calibration pilot measures chi_B at
 * calib_probes midpoints of [j_perp_start, j_perp_end], and J_perp(step) inverts the cumulative
 * integral of chi_B^alpha (eps_tilde <= 0), or J_perp advances online by eps_tilde chi_B^-alpha
 * (eps_tilde > 0). Both modes: the B sums are exact integers, so chi_B and the schedule follow the
 * reference's arithmetic; FP32 mode sweeps as qa_sqa_run does. debug_csv_path (NULL or "": none)
 * writes the reference's per-step diagnostic CSV. */

/* Replaces SQAAnnealer::run_optimal (sqa_annealer.cpp:499-952): a beta ramp at j_perp_start
 * for beta_ramp_fraction of the steps, then the adaptive J_perp schedule at beta. */
typedef struct qa_sqa_opt_params {
    double beta, j_perp_start, j_perp_end, eps_tilde, alpha;
    int64_t num_steps, sweeps_per_step, worldline_sweeps, cluster_sweeps, calib_probes, calib_sweeps;
    double beta_ramp_fraction, beta_ramp_start;
    int64_t trotter_slices, replicas;
    int32_t mode;
    uint64_t seed;                 /* FP32: Philox key                                         */
    const int8_t *init_spins;      /* validation: [replicas*P*n], SQAState::random draws       */
    const uint64_t *replica_seeds; /* validation: [replicas]                                   */
    const char *debug_csv_path;
} qa_sqa_opt_params;

typedef struct qa_opt_schedule {
    double *j_perp_trace;          /* [num_steps]                                               */
    double *chi_B_trace;           /* [num_steps] chi_B that drove each step (floored)          */
    double *calibrated_eps_tilde, *j_perp_start, *resolved_j_perp_end, *final_j_perp;
} qa_opt_schedule;

typedef struct qa_sqa_opt_result {
    int8_t *best_state;            /* [n]                                                       */
    double *best_energy;           /* FP64 energy of best_state                                 */
    double *energy_trace;          /* [num_steps] mean slice energy at each step end            */
    qa_opt_schedule sched;
    double *kernel_ms;             /* optional                                                  */
} qa_sqa_opt_result;

qa_status qa_sqa_optimal_run(qa_ctx *ctx, const qa_problem *p, const qa_sqa_opt_params *prm,
                             qa_sqa_opt_result *res);

/* Replaces SQAParallelTemperingAnnealer::run_optimal (sqa_parallel_tempering.cpp:348-882): every
 * rung sweeps at its beta/P with one shared J_perp from trotter_coupling(betas[0], gammas[0]) to
 * j_perp_end (<= 0: the coldest rung's coupling); swaps compare (beta_r/P) sum_t E_t. */
typedef struct qa_sqapt_opt_params {
    int64_t rungs;
    const double *betas, *gammas;  /* [rungs]                                                   */
    int64_t trotter_slices;        /* P actually used (caller resolves continuous_time_slices)  */
    int64_t num_steps, sweeps_per_step, worldline_sweeps, cluster_sweeps, swap_interval;
    int64_t calib_probes, calib_sweeps;
    double eps_tilde, alpha, j_perp_end;
    int32_t mode;
    uint64_t seed;                 /* FP32: Philox key (the pilot uses a key derived from it)   */
    const int8_t *init_spins;      /* validation: [rungs*P*n]                                   */
    const uint64_t *rung_seeds;    /* validation: [rungs]                                       */
    const uint64_t *scratch_seeds; /* validation, eps_tilde <= 0: [rungs] the pilot's engines   */
    const double *swap_uniforms;   /* validation: master-engine uniform stream for swaps        */
    int64_t swap_uniforms_len;
    const char *debug_csv_path;
} qa_sqapt_opt_params;

typedef struct qa_sqapt_opt_result {
    int8_t *final_states;          /* [rungs*n] best slice per rung at the end                  */
    double *final_energies;        /* [rungs]                                                   */
    int8_t *best_state;            /* [n]                                                       */
    double *best_energy;
    double *average_energy_trace;  /* [num_steps]                                               */
    double *swap_acceptance_trace; /* [num_steps]                                               */
    qa_opt_schedule sched;
    int64_t *swap_uniforms_used;   /* validation (may be NULL)                                  */
    double *kernel_ms;
} qa_sqapt_opt_result;

qa_status qa_sqapt_optimal_run(qa_ctx *ctx, const qa_problem *p, const qa_sqapt_opt_params *prm,
                               qa_sqapt_opt_result *res);

/* ---- sqa_chi: replaces SQAChiAnnealer::run_chi (sqa_chi_annealer.cpp:146-438) ------------ */
typedef struct qa_chi_params {
    int64_t trotter_slices;
    int64_t replicas;
    double beta, gamma_start, gamma_end;
    int64_t num_steps, sweeps_per_step, scan_points, scan_sweeps, scan_burn;
    double chi_floor_fraction, driver_A0, beta_ramp_fraction, beta_ramp_start;
    int32_t mode;
    uint64_t seed;
    /* validation: the master-engine state is consumed by the library; the caller passes the
     * seeds it would draw: [scan_points+1] fresh worldlines of R*P*n init spins and R*P cell
     * seeds each, in draw order. */
    const int8_t *init_spins;    /* [(scan_points+1)*R*P*n]                                   */
    const uint64_t *cell_seeds;  /* [(scan_points+1)*R*P]                                     */
    /* FP32 mode: when non-NULL the pilot scan is skipped and these chi_B values are used. */
    const double *chi_override;  /* [scan_points]                                            */
    /* FP32 mode, chi_override NULL: when non-NULL the chi_B profile is predicted by this trained
     * surrogate (n_out == scan_points) instead of scanned; the pilot scan still runs when the
     * ensemble spread exceeds surrogate_max_spread or the instance lies outside the training
     * range of the surrogate's features. */
    const struct qa_mlp *surrogate;
    double surrogate_max_spread;
} qa_chi_params;

typedef struct qa_chi_result {
    int8_t *best_state;
    double *best_energy;
    double *energy_trace, *beta_schedule, *gamma_schedule, *s_schedule;   /* [num_steps]    */
    double *scan_s, *scan_gamma, *scan_chi_B;                             /* [scan_points]  */
    double *scalars;             /* [s_star, gamma_star, j_perp_star, chi_floor, driver_A0]  */
    double *kernel_ms;
    int32_t *surrogate_used;     /* optional: 1 if the surrogate's profile was used, 0 if the
                                    pilot ran (no surrogate, or the fallback was taken)        */
    double *surrogate_spread;    /* optional: ensemble spread of the prediction (+inf when the
                                    features were outside the training range)                */
} qa_chi_result;

qa_status qa_sqa_chi_run(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm,
                         qa_chi_result *res);

/* Stages 1-2 of run_chi without the production anneal (sqa_chi_annealer.cpp:146-330), FP32 mode:
 * the chi_B profile (chi_override, else the surrogate, else the pilot scan, exactly as
 * qa_sqa_chi_run chooses it) and the beta/gamma/s schedules built from it, bit-identical to the
 * schedules qa_sqa_chi_run returns for the same prm. Fills beta_schedule, gamma_schedule,
 * s_schedule, scan_s, scan_gamma, scan_chi_B, scalars and the optional kernel_ms (pilot time),
 * surrogate_used and surrogate_spread; best_state, best_energy and energy_trace are not used and
 * may be NULL. For callers that anneal on that schedule themselves (an SQA-PT ladder per step,
 * qa_sqapt_params.step_betas / step_gammas). */
qa_status qa_chi_schedule(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, qa_chi_result *res);

/* Stage 1 of run_chi alone (the pilot chi_B scan, sqa_chi_annealer.cpp:232-281), FP32 mode: the
 * profile the production schedule is built from, and the training label of the surrogate.
 * Uses beta, gamma_start, gamma_end, driver_A0, trotter_slices, replicas, scan_points,
 * scan_sweeps, scan_burn and seed of prm. */
qa_status qa_chi_pilot(qa_ctx *ctx, const qa_problem *p, const qa_chi_params *prm, double *scan_s,
                       double *scan_gamma, double *scan_chi_B, double *kernel_ms);

/* ---- chi_B surrogate: no counterpart in the reference ----------------------------------------
 * For callers of run_chi that want the schedule without the pilot scan. Features of an instance
 * (QA_CHI_NFEATURES doubles): log2 n, log2(1 + couplings per spin), log field scale, field-scale
 * heterogeneity, log coupling kurtosis, negative-coupling fraction, mean |h| / field scale, mean
 * interaction order, log(beta * field scale), log P, log(A0 / field scale),
 * log(gamma_end / gamma_start). The network is an ensemble of `members` MLPs
 * n_in -> width (tanh) x depth -> n_out (sigmoid) trained on normalized pilot profiles
 * chi_B / max chi_B; its ensemble mean is the predicted profile and the largest per-point
 * standard deviation across members is its spread. */
#define QA_CHI_NFEATURES 12
qa_status qa_chi_features(const qa_problem *p, double beta, int64_t trotter_slices, double driver_A0,
                          double gamma_start, double gamma_end, double *features);

typedef struct qa_mlp qa_mlp;
typedef struct qa_mlp_train_params {
    int64_t epochs, batch;
    double lr, beta1, beta2, eps, weight_decay;  /* Adam; 0 picks 0.9 / 0.999 / 1e-8         */
    uint64_t seed;                               /* minibatch order                           */
} qa_mlp_train_params;

qa_status qa_mlp_create(qa_ctx *ctx, int32_t n_in, int32_t width, int32_t depth, int32_t n_out,
                        int32_t members, uint64_t seed, qa_mlp **out);
qa_status qa_mlp_destroy(qa_mlp *m);
/* parameters per member; layout of one member: for each layer W [out][in] then b [out] */
qa_status qa_mlp_info(const qa_mlp *m, int64_t *params_per_member, int32_t *trained);
/* params [members * params_per_member]; norm [4 * n_in]: mean, scale, training low, high */
qa_status qa_mlp_get_state(const qa_mlp *m, float *params, double *norm);
qa_status qa_mlp_set_state(qa_mlp *m, const float *params, const double *norm);
/* y [members][batch][n_out] from raw features x [batch][n_in] */
qa_status qa_mlp_forward(qa_mlp *m, int64_t batch, const double *x, float *y);
/* mean squared error per member against t [batch][n_out] and its gradient [members][params] */
qa_status qa_mlp_loss_grad(qa_mlp *m, int64_t batch, const double *x, const float *t, double *loss,
                           float *grad);
/* Adam on minibatches; sets the input normalization from x; mean loss per epoch */
qa_status qa_mlp_train(qa_mlp *m, int64_t samples, const double *x, const float *t,
                       const qa_mlp_train_params *tp, double *loss_history);

#ifdef __cplusplus
}
#endif
#endif
