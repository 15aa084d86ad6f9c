"""Adaptive J_perp schedules (run_optimal) through the drop-in classes on the GPU.

FP64 validation ("hip_fp64"): SQAAnnealer::run_optimal and SQAParallelTemperingAnnealer::run_optimal
bit-equal to the golden fixtures written from the reference build (sqa_annealer.cpp:499-952,
sqa_parallel_tempering.cpp:348-882): best state and energy, energy traces, the J_perp and chi_B
traces and the calibration scalars. FP32 ("hip"): returned energies equal the FP64 energy of the
returned states, the schedule keeps its construction invariants, and best-energy distributions
over seeds agree with the oracle's (two-sample KS, p > 0.01, energies rounded to 1e-9).
"""
import numpy as np
import pytest

from tests.golden_io import load, out, dec, bits_equal, problem

pytestmark = pytest.mark.gpu

qa = pytest.importorskip("qlatanneal_amd")


def _model(d):
    p = problem(d)
    if p.kind == 0:
        return p, qa.DenseIsing(p.h, p.J, p.c)
    if p.kind == 1:
        return p, qa.SparseIsing(p.h, [qa.SparseEdge(i, j, v) for i, j, v in p.edges], p.n, p.c)
    return p, qa.HigherOrderIsing([(list(v), c) for v, c in p.terms] + [([], p.c)], p.n)


def _sk(rng, n):
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def _sched_fields(r):
    return dict(j_perp_trace=np.array(r.j_perp_trace), chi_B_trace=np.array(r.chi_B_trace),
                calibrated_eps_tilde=np.array([r.calibrated_eps_tilde]), j_perp_start=np.array([r.j_perp_start]),
                resolved_j_perp_end=np.array([r.resolved_j_perp_end]), final_j_perp=np.array([r.final_j_perp]))


def test_sqa_optimal_validation_golden():
    for c in load("sqa_opt.json")["cases"]:
        print("sqa_opt case seed", c["seed"], flush=True)
        _, m = _model(c["problem"])
        a = qa.SQAAnnealer(m, qa.SQASchedule.from_vectors([c["beta"]], [1.0]), c["trotter_slices"], c["replicas"],
                           backend="hip_fp64")
        a.set_seed(c["seed"])
        r = a.run_optimal(c["beta"], c["j_perp_start"], c["j_perp_end"], c["eps_tilde"], c["alpha"], c["num_steps"],
                          c["sweeps_per_step"], c["worldline_sweeps"], c["cluster_sweeps"], c["calib_probes"],
                          c["calib_sweeps"], "", c["beta_ramp_fraction"], c["beta_ramp_start"])
        got = dict(best_state=np.array(r.best_state.spins, np.int8), best_energy=np.array([r.best_energy]),
                   energy_trace=np.array(r.energy_trace), **_sched_fields(r))
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(np.ravel(got[k]), np.ravel(exp[k])), (c["seed"], k)


def test_sqapt_optimal_validation_golden():
    for c in load("sqapt_opt.json")["cases"]:
        print("sqapt_opt case seed", c["seed"], flush=True)
        _, m = _model(c["problem"])
        a = qa.SQAParallelTemperingAnnealer(m, dec(c["betas"]).tolist(), dec(c["gammas"]).tolist(),
                                            c["trotter_slices"], backend="hip_fp64")
        a.set_seed(c["seed"])
        r = a.run_optimal(c["num_steps"], c["sweeps_per_step"], c["worldline_sweeps"], c["eps_tilde"], c["alpha"],
                          c["j_perp_end"], c["cluster_sweeps"], c["swap_interval"], 0, c["calib_probes"],
                          c["calib_sweeps"])
        got = dict(final_states=np.array([list(s.spins) for s in r.final_states], np.int8),
                   final_energies=np.array(r.final_energies), best_state=np.array(r.best_state.spins, np.int8),
                   best_energy=np.array([r.best_energy]), average_energy_trace=np.array(r.average_energy_trace),
                   swap_acceptance_trace=np.array(r.swap_acceptance_trace), **_sched_fields(r))
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(np.ravel(got[k]), np.ravel(exp[k])), (c["seed"], k)


def test_sqa_optimal_debug_csv(tmp_path):
    """The diagnostic CSV has the reference's header and one row per probe and per step."""
    c = load("sqa_opt.json")["cases"][0]
    _, m = _model(c["problem"])
    a = qa.SQAAnnealer(m, qa.SQASchedule.from_vectors([1.0], [1.0]), c["trotter_slices"], c["replicas"],
                       backend="hip_fp64")
    a.set_seed(c["seed"])
    path = tmp_path / "opt.csv"
    a.run_optimal(c["beta"], c["j_perp_start"], c["j_perp_end"], c["eps_tilde"], c["alpha"], c["num_steps"],
                  c["sweeps_per_step"], c["worldline_sweeps"], c["cluster_sweeps"], c["calib_probes"],
                  c["calib_sweeps"], str(path), c["beta_ramp_fraction"], c["beta_ramp_start"])
    rows = path.read_text().splitlines()
    assert rows[0] == "phase,step_index,j_perp,chi_B,delta_j,eps_tilde,alpha,floor_hit,chi_B_online"
    phases = [r.split(",")[0] for r in rows[1:]]
    n1 = int(c["beta_ramp_fraction"] * c["num_steps"])
    assert phases.count("beta_ramp") == n1 and phases.count("calib") == c["calib_probes"]
    assert phases.count("run") == c["num_steps"] - n1


def test_sqa_optimal_fp32_invariants_and_distribution():
    from scipy.stats import ks_2samp
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(41)
    n, P, R, steps = 32, 8, 4, 24
    J = _sk(rng, n)
    m = qa.DenseIsing(np.zeros(n), J)
    a = qa.SQAAnnealer(m, qa.SQASchedule.from_vectors([1.0], [1.0]), P, R, backend="hip")
    a.set_seed(3)
    r = a.run_optimal(4.0, 0.05, 3.0, 0.0, 15 / 14, steps, 2, 1, 0, 6, 3, "", 0.25, 0.1)
    jt = np.array(r.j_perp_trace)
    n1 = int(0.25 * steps)
    assert len(r.energy_trace) == steps and len(jt) == steps and len(r.chi_B_trace) == steps
    assert np.all(jt[:n1] == 0.05) and jt[n1] == 0.05 and abs(jt[-1] - 3.0) < 1e-12
    assert np.all(np.diff(jt[n1:]) >= 0) and r.calibrated_eps_tilde > 0
    assert abs(m.energy(list(r.best_state.spins)) - r.best_energy) < 1e-9
    # online schedule (eps_tilde > 0) on the GPU: never passes j_perp_end, single replica guard
    b = qa.SQAAnnealer(m, qa.SQASchedule.from_vectors([1.0], [1.0]), P, 1, backend="hip")
    b.set_seed(5)
    o = b.run_optimal(4.0, 0.05, 3.0, 0.02, 15 / 14, 12, 1, 0, 1, 6, 3, "", 0.0, 0.1)
    assert np.all(np.array(o.j_perp_trace) <= 3.0) and np.all(np.diff(o.j_perp_trace) >= 0)
    assert abs(m.energy(list(o.best_state.spins)) - o.best_energy) < 1e-9
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    p = Problem.dense(np.zeros(n), J)
    gpu, cpu = [], []
    for k in range(64):
        a.set_seed(100 + k)
        gpu.append(a.run_optimal(4.0, 0.05, 3.0, 0.0, 15 / 14, 16, 1, 0, 0, 4, 2, "", 0.25, 0.1).best_energy)
        cpu.append(orc.sqa_opt_run(p, P, R, 4.0, 0.05, 3.0, 0.0, 15 / 14, 16, 1, 0, 0, 4, 2, 0.25, 0.1,
                                   seed=900 + k)["best_energy"])
    pv = ks_2samp(np.round(gpu, 9), np.round(cpu, 9)).pvalue
    print("sqa_opt KS p", pv, "gpu mean", np.mean(gpu), "cpu mean", np.mean(cpu), flush=True)
    assert pv > 0.01


def test_sqapt_optimal_fp32_invariants_and_distribution():
    from scipy.stats import ks_2samp
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(42)
    n, P = 24, 6
    J = _sk(rng, n)
    m = qa.DenseIsing(np.zeros(n), J)
    betas, gammas = np.geomspace(0.5, 4.0, 4).tolist(), np.geomspace(2.5, 0.5, 4).tolist()
    a = qa.SQAParallelTemperingAnnealer(m, betas, gammas, P, backend="hip")
    a.set_seed(7)
    r = a.run_optimal(16, 2, 1, 0.0, 15 / 14, 3.0, 0, 2, 0, 5, 2)
    for s, e in zip(r.final_states, r.final_energies):
        assert abs(m.energy(list(s.spins)) - e) < 1e-9
    assert abs(m.energy(list(r.best_state.spins)) - r.best_energy) < 1e-9
    assert r.best_energy <= min(r.final_energies) + 1e-4  # best tracked on FP32-field energies
    jt = np.array(r.j_perp_trace)
    assert jt[0] == r.j_perp_start and abs(jt[-1] - 3.0) < 1e-12 and np.all(np.diff(jt) >= 0)
    sw = np.array(r.swap_acceptance_trace)
    assert np.all(sw[0::2] == 0.0) and np.all((sw >= 0) & (sw <= 1))
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    p = Problem.dense(np.zeros(n), J)
    gpu, cpu = [], []
    for k in range(64):
        a.set_seed(200 + k)
        gpu.append(a.run_optimal(10, 1, 0, 0.0, 15 / 14, 3.0, 0, 1, 0, 4, 2).best_energy)
        cpu.append(orc.sqapt_opt_run(p, betas, gammas, P, 10, 1, 0, 0.0, 15 / 14, 3.0, 0, 1, 0, 4, 2,
                                     seed=700 + k)["best_energy"])
    pv = ks_2samp(np.round(gpu, 9), np.round(cpu, 9)).pvalue
    print("sqapt_opt KS p", pv, "gpu mean", np.mean(gpu), "cpu mean", np.mean(cpu), flush=True)
    assert pv > 0.01


def test_solve_optimal_schedules():
    rng = np.random.default_rng(43)
    n = 20
    J = _sk(rng, n)
    d = qa.DenseIsing(np.zeros(n), J)
    for method in ("sqa", "sqapt"):
        res = qa.solve(d, method=method, schedule_type="optimal", reads=2, seed=4)
        assert len(res.energies) == 2
        for s, e in zip(res.samples, res.energies):
            assert abs(d.energy(np.asarray(s).astype(int).tolist()) - e) < 1e-9
