/*
 * Oracle C-ABI — TEST INFRASTRUCTURE ONLY.
 *
 * Two libraries implement this header:
 *   oracle/liboracle.so         my CPU restatement of the reference engines (orc_restated.cpp)
 *   oracle/_ref/libqanneal_ref.so  the reference sources compiled by oracle/Makefile plus a shim
 *                                  (orc_ref_shim.cpp) that calls the reference classes.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load either library,
 * and only as the checker or the timed CPU baseline; the product never links it.
 *
 * All arrays are caller-allocated. Functions return 0 on success and nonzero on error
 * (orc_last_error() then holds a message).
 */
#ifndef ORC_ABI_H
#define ORC_ABI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_DENSE = 0, ORC_SPARSE = 1, ORC_HUBO = 2 };

typedef struct orc_problem {
    int32_t kind;              /* ORC_DENSE / ORC_SPARSE / ORC_HUBO */
    int64_t n;
    const double *h;           /* dense, sparse: [n] */
    const double *J;           /* dense: [n*n] row-major */
    double c;                  /* constant offset */
    int64_t nedges;            /* sparse: undirected edge list in caller order */
    const int64_t *ei;
    const int64_t *ej;
    const double *ev;
    int64_t nterms;            /* hubo: raw terms, CSR over term_vars (any arity, repeats allowed) */
    const int64_t *term_off;   /* [nterms+1] */
    const int64_t *term_vars;
    const double *term_coeff;  /* [nterms] */
} orc_problem;

const char *orc_last_error(void);
int orc_set_threads(int threads);

/* Hamiltonian contract (hamiltonian.hpp:10-50). spins: count*n int8 in {-1,+1}. */
int orc_energy(const orc_problem *p, int64_t count, const int8_t *spins, double *out);
int orc_delta_energy(const orc_problem *p, const int8_t *spins, int64_t flip, double *out);
int orc_local_fields(const orc_problem *p, const int8_t *spins, double *fields);
int orc_update_fields(const orc_problem *p, double *fields, const int8_t *new_spins,
                      int64_t flip, int8_t old_spin);
/* QUBO::to_ising (qubo.cpp:33-69): Q [n*n] -> h [n], J [n*n], c */
int orc_qubo_to_ising(int64_t n, const double *Q, double *h, double *J, double *c);

/* Annealer::run (annealer.cpp:33-75) after set_seed(seed). trace: [steps] */
int orc_sa_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
               uint64_t seed, int8_t *best_state, double *best_energy, double *trace);

/* ReplicaAnnealer::run (replica_annealer.cpp:46-142) after set_seed(seed). Per replica: best
 * state/energy, energy and magnetization traces [replicas*steps]; global best; average traces
 * [steps]. */
int orc_replica_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
                    int64_t replicas, uint64_t seed, int8_t *best_states, double *best_energies,
                    double *energy_traces, double *mag_traces, int8_t *global_best_state,
                    double *global_best_energy, double *avg_energy_trace, double *avg_mag_trace);

/* ParallelTemperingAnnealer::run (parallel_tempering.cpp:36-122) after set_seed(seed).
 * final_states [replicas*n], final_energies [replicas], avg_trace/swap_trace [steps]. */
int orc_pt_run(const orc_problem *p, int64_t replicas, const double *betas, int64_t sweeps_per_step,
               int64_t steps, int64_t swap_interval, uint64_t seed, int8_t *final_states,
               double *final_energies, int8_t *best_state, double *best_energy, double *avg_trace,
               double *swap_trace);

/* SQAAnnealer::run (sqa_annealer.cpp:87-477) after set_seed(seed). trace: [steps] */
int orc_sqa_run(const orc_problem *p, int64_t steps, const double *betas, const double *gammas,
                int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                int64_t worldline_sweeps, int64_t cluster_sweeps, int64_t continuous_time_slices,
                uint64_t seed, int8_t *best_state, double *best_energy, double *trace);

/* set_seed(seed), then `runs` consecutive SQAAnnealer::run calls on one annealer (the engine
 * continues from run to run); best_state/best_energy/trace of the last run. */
int orc_sqa_run_repeat(const orc_problem *p, int64_t runs, int64_t steps, const double *betas,
                       const double *gammas, int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                       int64_t worldline_sweeps, int64_t cluster_sweeps, int64_t continuous_time_slices,
                       uint64_t seed, int8_t *best_state, double *best_energy, double *trace);

/* SQAParallelTemperingAnnealer::run (sqa_parallel_tempering.cpp:81-346) after set_seed(seed).
 * final_states [rungs*n], final_energies [rungs], avg_trace/swap_trace [steps]. */
int orc_sqapt_run(const orc_problem *p, int64_t rungs, const double *betas, const double *gammas,
                  int64_t trotter_slices, int64_t sweeps_per_step, int64_t worldline_sweeps,
                  int64_t steps, int64_t swap_interval, int64_t cluster_sweeps,
                  int64_t continuous_time_slices, uint64_t seed, int8_t *final_states,
                  double *final_energies, int8_t *best_state, double *best_energy,
                  double *avg_trace, double *swap_trace);

/* Annealer::run (annealer.cpp:33-75) after set_seed(seed) with one observer that is an Observer
 * and a SweepObserver: step_mag [steps] the mean spin after each step; sweep_energies
 * [steps*sweeps_per_beta] and sweep_states [steps*sweeps_per_beta*n] after each sweep. */
int orc_sa_observe_run(const orc_problem *p, int64_t steps, const double *betas, int64_t sweeps_per_beta,
                       uint64_t seed, int8_t *best_state, double *best_energy, double *trace, double *step_mag,
                       double *sweep_energies, int8_t *sweep_states);

/* SQAAnnealer::run (sqa_annealer.cpp:87-477) after set_seed(seed) with an SQAObserver
 * (tracing = 0: the parallel path) or an observer that is also an SQASweepObserver (tracing = 1:
 * the sequential path, :317-453). step_mag [steps]: mean spin of the lattice after each step
 * (SQAMetricsObserver). tracing = 1, K = steps*replicas*(sweeps_per_beta+cluster+worldline)
 * records: sweep_meta [K*4] (step, replica, sweep, phase), sweep_avg [K], sweep_rep [K*replicas],
 * sweep_states [K*replicas*P*n]. */
int orc_sqa_observe_run(const orc_problem *p, int64_t steps, const double *betas, const double *gammas,
                        int64_t trotter_slices, int64_t replicas, int64_t sweeps_per_beta,
                        int64_t worldline_sweeps, int64_t cluster_sweeps, int32_t tracing, uint64_t seed,
                        int8_t *best_state, double *best_energy, double *trace, double *step_mag,
                        int64_t *sweep_meta, double *sweep_avg, double *sweep_rep, int8_t *sweep_states);

/* SQAAnnealer::run_optimal (sqa_annealer.cpp:499-952) after set_seed(seed). energy_trace,
 * j_perp_trace, chi_trace: [num_steps]; scalars [4]: calibrated_eps_tilde, j_perp_start,
 * resolved_j_perp_end, final_j_perp. No debug CSV. */
int orc_sqa_opt_run(const orc_problem *p, int64_t trotter_slices, int64_t replicas, double beta,
                    double j_perp_start, double j_perp_end, double eps_tilde, double alpha,
                    int64_t num_steps, int64_t sweeps_per_step, int64_t worldline_sweeps,
                    int64_t cluster_sweeps, int64_t calib_probes, int64_t calib_sweeps,
                    double beta_ramp_fraction, double beta_ramp_start, uint64_t seed,
                    int8_t *best_state, double *best_energy, double *energy_trace,
                    double *j_perp_trace, double *chi_trace, double *scalars);

/* SQAParallelTemperingAnnealer::run_optimal (sqa_parallel_tempering.cpp:348-882) after
 * set_seed(seed). final_states [rungs*n], final_energies [rungs], the traces [num_steps],
 * scalars [4] as above. No debug CSV. */
int orc_sqapt_opt_run(const orc_problem *p, int64_t rungs, const double *betas, const double *gammas,
                      int64_t trotter_slices, int64_t num_steps, int64_t sweeps_per_step,
                      int64_t worldline_sweeps, double eps_tilde, double alpha, double j_perp_end,
                      int64_t cluster_sweeps, int64_t swap_interval, int64_t continuous_time_slices,
                      int64_t calib_probes, int64_t calib_sweeps, uint64_t seed, int8_t *final_states,
                      double *final_energies, int8_t *best_state, double *best_energy,
                      double *avg_trace, double *swap_trace, double *j_perp_trace, double *chi_trace,
                      double *scalars);

/* SQAChiAnnealer::run_chi (sqa_chi_annealer.cpp:146-438) after set_seed(seed).
 * trace/beta_sched/gamma_sched/s_sched: [num_steps]; scan_*: [scan_points];
 * scalars: [s_star, gamma_star, j_perp_star, chi_floor, driver_A0]. */
int orc_sqa_chi_run(const orc_problem *p, int64_t trotter_slices, int64_t replicas, double beta,
                    double gamma_start, double gamma_end, int64_t num_steps,
                    int64_t sweeps_per_step, int64_t scan_points, int64_t scan_sweeps,
                    int64_t scan_burn, double chi_floor_fraction, double driver_A0,
                    double beta_ramp_fraction, double beta_ramp_start, uint64_t seed,
                    int8_t *best_state, double *best_energy, double *trace, double *beta_sched,
                    double *gamma_sched, double *s_sched, double *scan_s, double *scan_gamma,
                    double *scan_chi, double *scalars);

#ifdef __cplusplus
}
#endif
#endif
