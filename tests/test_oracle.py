"""Pins the CPU oracle (oracle/liboracle.so) before anything is checked against it.

1. Bit-exact agreement with the golden fixtures written from the reference build.
2. Checks fixed by the mathematics alone, chosen so that a dropped term, a wrong sign or a
   transposed operand in the restatement fails: brute-force energy definitions, dE against
   E(flipped)-E, incremental fields against recomputed ones, QUBO<->Ising equality on every
   bit string, and Metropolis/Trotter stationary means against exact enumeration.
3. Where the reference build is present (only in the container that holds the reference
   tree), a randomized oracle-vs-reference sweep over all engines.
"""
import itertools
import os

import numpy as np
import pytest

from oracle.orc import Oracle, Problem, ORACLE_SO, REF_SO
from tests.golden_io import load, problem, out, dec, bits_equal


@pytest.fixture(scope="module")
def orc():
    o = Oracle(ORACLE_SO)
    o.set_threads(1)
    return o


# ---------------------------------------------------------------------------------------------
# 1. golden fixtures
# ---------------------------------------------------------------------------------------------
def test_golden_hamiltonian(orc):
    g = load("hamiltonian.json")
    for c in g["cases"]:
        p = problem(c["problem"])
        s = np.array(c["spins"], np.int8)
        assert bits_equal(orc.energy(p, s), dec(c["energy"]))
        assert bits_equal([orc.delta_energy(p, s[0], c["flip"])], dec(c["delta"]))
        f = orc.local_fields(p, s[0])
        assert bits_equal(f, dec(c["fields"]))
        ns = s[0].copy()
        old = int(ns[c["flip"]])
        ns[c["flip"]] = -old
        assert bits_equal(orc.update_fields(p, f, ns, c["flip"], old), dec(c["updated_fields"]))
    # tests/test_dense_ising.cpp:21-22 known answer
    assert dec(g["cases"][0]["energy"])[0] == 0.5
    for q in g["qubo"]:
        n = q["n"]
        h, J, c = orc.qubo_to_ising(dec(q["Q"]).reshape(n, n))
        assert bits_equal(h, dec(q["h"])) and bits_equal(J.ravel(), dec(q["J"])) and bits_equal([c], dec(q["c"]))


def test_golden_sa(orc):
    for c in load("sa.json")["cases"]:
        res = orc.sa_run(problem(c["problem"]), dec(c["betas"]), c["sweeps_per_beta"], c["seed"])
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(res[k], exp[k]), k


def test_golden_replica(orc):
    for c in load("replica.json")["cases"]:
        res = orc.replica_run(problem(c["problem"]), dec(c["betas"]), c["sweeps_per_beta"], c["replicas"], c["seed"])
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(np.ravel(res[k]), np.ravel(exp[k])), k


def test_golden_pt(orc):
    for c in load("pt.json")["cases"]:
        res = orc.pt_run(problem(c["problem"]), dec(c["betas"]), c["sweeps_per_step"], c["steps"],
                         c["swap_interval"], c["seed"])
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(np.ravel(res[k]), np.ravel(exp[k])), k


def test_replica_pt_closed_forms(orc):
    """ReplicaAnnealer: the average traces are the sequential means of the per-replica traces and the
    global best is the smallest replica best. ParallelTempering: with every beta equal a swap has
    dE-weight 0 and is always accepted (trace exactly 1), and with swaps every step the best energy
    is the exact energy of the best state."""
    rng = np.random.default_rng(12)
    n = 14
    J = rng.normal(size=(n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    p = Problem.dense(rng.normal(size=n) * 0.2, J, 0.3)
    r = orc.replica_run(p, np.linspace(0.2, 2.0, 9), 2, 5, 44)
    for k, key in (("energy_traces", "average_energy_trace"), ("magnetization_traces", "average_magnetization_trace")):
        acc = np.zeros(9)
        for row in r[k]:
            acc = acc + row
        assert bits_equal(acc / 5.0, r[key])
    assert r["global_best_energy"] == r["best_energies"].min()
    assert np.all(np.abs(r["magnetization_traces"]) <= 1.0)
    t = orc.pt_run(p, np.full(4, 1.3), 1, 12, 1, 9)
    assert np.all(t["swap_acceptance_trace"] == 1.0)
    assert abs(orc.energy(p, t["best_state"])[0] - t["best_energy"]) < 1e-9
    assert np.allclose(orc.energy(p, t["final_states"]), t["final_energies"], atol=1e-9)


def _opt_args(c, keys):
    return {k: c[k] for k in keys}


SQA_OPT_KEYS = ("trotter_slices", "replicas", "beta", "j_perp_start", "j_perp_end", "eps_tilde", "alpha",
                "num_steps", "sweeps_per_step", "worldline_sweeps", "cluster_sweeps", "calib_probes", "calib_sweeps",
                "beta_ramp_fraction", "beta_ramp_start", "seed")
SQAPT_OPT_KEYS = ("trotter_slices", "num_steps", "sweeps_per_step", "worldline_sweeps", "eps_tilde", "alpha",
                  "j_perp_end", "cluster_sweeps", "swap_interval", "calib_probes", "calib_sweeps", "seed")


def test_golden_sqa_optimal(orc):
    for c in load("sqa_opt.json")["cases"]:
        res = orc.sqa_opt_run(problem(c["problem"]), **_opt_args(c, SQA_OPT_KEYS))
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(np.ravel(res[k]), np.ravel(exp[k])), k


def test_golden_sqapt_optimal(orc):
    for c in load("sqapt_opt.json")["cases"]:
        res = orc.sqapt_opt_run(problem(c["problem"]), dec(c["betas"]), dec(c["gammas"]),
                                **_opt_args(c, SQAPT_OPT_KEYS))
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(np.ravel(res[k]), np.ravel(exp[k])), k


def test_optimal_schedule_invariants(orc):
    """What the schedule construction fixes (sqa_annealer.cpp:777-949): the beta-ramp steps sit at
    j_perp_start; the calibrated trajectory starts exactly at j_perp_start, is non-decreasing and ends
    at j_perp_end to rounding; eps_tilde equals the probe integral over the phase-2 budget; every
    chi_B entry is at least the floor; the online schedule never passes j_perp_end; best_energy is
    the FP64 energy of best_state; and the SQA-PT start J_perp is the hottest rung's Trotter coupling."""
    rng = np.random.default_rng(3)
    n = 12
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    p = Problem.dense(np.zeros(n), J)
    r = orc.sqa_opt_run(p, 6, 2, 3.0, 0.05, 2.5, 0.0, 15 / 14, 20, 2, 0, 0, 6, 3, 0.3, 0.1, 5)
    jt = r["j_perp_trace"]
    n1 = 6
    assert np.all(jt[:n1] == 0.05)
    assert jt[n1] == 0.05 and np.all(np.diff(jt[n1:]) >= 0) and abs(jt[-1] - 2.5) < 1e-12
    assert r["final_j_perp"] == jt[-1]
    assert np.all(r["chi_B_trace"][n1:] > 0)
    assert abs(orc.energy(p, r["best_state"])[0] - r["best_energy"]) < 1e-12
    o = orc.sqa_opt_run(p, 6, 2, 3.0, 0.05, 2.5, 0.02, 15 / 14, 20, 2, 0, 0, 6, 3, 0.0, 0.1, 5)
    assert o["calibrated_eps_tilde"] == 0.02 and np.all(np.diff(o["j_perp_trace"]) >= 0)
    assert np.all(o["j_perp_trace"] <= 2.5) and o["j_perp_trace"][0] == 0.05
    betas, gammas = np.array([0.5, 1.5, 3.0]), np.array([2.0, 1.0, 0.5])
    t = orc.sqapt_opt_run(p, betas, gammas, 4, 10, 1, 0, 0.0, 15 / 14, 2.0, 0, 1, 0, 4, 2, 9)
    jp0 = 0.5 * np.log(1.0 / np.tanh(max(0.5 * 2.0 / 4, 1e-12)))
    assert abs(t["j_perp_start"] - jp0) < 1e-15 and t["j_perp_trace"][0] == t["j_perp_start"]
    assert abs(t["j_perp_trace"][-1] - 2.0) < 1e-12
    d = orc.sqapt_opt_run(p, betas, gammas, 4, 6, 1, 0, 0.0, 15 / 14, 0.0, 0, 1, 0, 4, 2, 9)
    cold = 0.5 * np.log(1.0 / np.tanh(max(3.0 * 0.5 / 4, 1e-12)))
    assert abs(d["resolved_j_perp_end"] - cold) < 1e-15 or d["resolved_j_perp_end"] == d["j_perp_start"] + 1e6


def test_golden_observers(orc):
    obs = load("observers.json")
    for c in obs["sa"]:
        res = orc.sa_observe_run(problem(c["problem"]), dec(c["betas"]), c["sweeps_per_beta"], c["seed"])
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(np.ravel(res[k]).astype(exp[k].dtype), np.ravel(exp[k])), k
    for c in obs["sqa"]:
        res = orc.sqa_observe_run(problem(c["problem"]), dec(c["betas"]), dec(c["gammas"]), c["trotter_slices"],
                                  c["replicas"], c["sweeps_per_beta"], c["worldline_sweeps"], c["cluster_sweeps"],
                                  c["tracing"], c["seed"])
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(np.ravel(res[k]).astype(exp[k].dtype), np.ravel(exp[k])), k


def test_observer_records_consistent(orc):
    """An observer does not change the SA trajectory (annealer.cpp:51-70): the result equals the
    unobserved run, the last sweep of each step is the step-end state, and the per-step energy
    is the running energy; the SQA tracing path's step tail is the mean over the final lattice."""
    rng = np.random.default_rng(8)
    n = 10
    J = rng.normal(size=(n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    p = Problem.dense(rng.normal(size=n) * 0.1, J)
    b = np.linspace(0.1, 2.0, 7)
    o = orc.sa_observe_run(p, b, 3, 5)
    r = orc.sa_run(p, b, 3, 5)
    assert bits_equal(o["energy_trace"], r["energy_trace"]) and o["best_energy"] == r["best_energy"]
    assert bits_equal(o["sweep_energies"][2::3], o["energy_trace"])
    last = o["sweep_states"][2::3]
    assert np.allclose(last.mean(axis=1), o["magnetization_trace"])
    t = orc.sqa_observe_run(p, b, np.geomspace(2, 0.1, 7), 3, 2, 1, 1, 1, True, 6)
    R, P = 2, 3
    per_step = R * 3
    final = t["sweep_states"][per_step - 1::per_step].reshape(7, R * P, n)
    e = np.array([[orc.energy(p, f[k:k + 1])[0] for k in range(R * P)] for f in final])
    assert np.allclose(e.mean(axis=1), t["energy_trace"], atol=1e-12)
    assert np.all(t["sweep_meta"][:, 3] == np.tile([0, 2, 1], 7 * R))


def test_golden_sqa(orc):
    for c in load("sqa.json")["cases"]:
        res = orc.sqa_run(problem(c["problem"]), dec(c["betas"]), dec(c["gammas"]), c["trotter_slices"],
                          c["replicas"], c["sweeps_per_beta"], c["worldline_sweeps"], c["cluster_sweeps"],
                          c["continuous_time_slices"], c["seed"])
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(res[k], exp[k]), k


def test_golden_sqa_repeat(orc):
    """Consecutive runs without reseeding continue the annealer's engine (sqa_annealer.cpp:100,
    108-111 draw from the member engine on every run): the oracle's last run equals the
    reference build's, a single repeat equals sqa_run, and the second run is not the first."""
    for c in load("sqa_repeat.json")["cases"]:
        p = problem(c["problem"])
        args = (dec(c["betas"]), dec(c["gammas"]), c["trotter_slices"], c["replicas"], c["sweeps_per_beta"],
                c["worldline_sweeps"], c["cluster_sweeps"], 0, c["seed"])
        res = orc.sqa_run_repeat(p, c["runs"], *args)
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(res[k], exp[k]), k
        one = orc.sqa_run(p, *args)
        first = orc.sqa_run_repeat(p, 1, *args)
        assert bits_equal(one["energy_trace"], first["energy_trace"])
        assert not bits_equal(one["energy_trace"], res["energy_trace"])


def test_golden_sqapt(orc):
    for c in load("sqapt.json")["cases"]:
        res = orc.sqapt_run(problem(c["problem"]), dec(c["betas"]), dec(c["gammas"]), c["trotter_slices"],
                            c["sweeps_per_step"], c["worldline_sweeps"], c["steps"], c["swap_interval"],
                            c["cluster_sweeps"], c["continuous_time_slices"], c["seed"])
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(res[k], exp[k]), k


def test_golden_sqa_chi(orc):
    for c in load("sqa_chi.json")["cases"]:
        res = orc.sqa_chi_run(problem(c["problem"]), c["trotter_slices"], c["replicas"], c["beta"],
                              c["gamma_start"], c["gamma_end"], c["num_steps"], c["sweeps_per_step"],
                              c["scan_points"], c["scan_sweeps"], c["scan_burn"], seed=c["seed"])
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(res[k], exp[k]), k


def test_golden_chi_exact_ground_states():
    # tests/test_sqa_chi_annealer.cpp:241-271: ferro chain -(n-1), sparse ferro ring -n
    cases = load("sqa_chi.json")["cases"]
    assert dec(cases[1]["out"]["best_energy"])[0] == -11.0
    assert dec(cases[2]["out"]["best_energy"])[0] == -12.0


# ---------------------------------------------------------------------------------------------
# 2. mathematics
# ---------------------------------------------------------------------------------------------
def _rand_problems(rng):
    n = 7
    J = rng.normal(size=(n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    yield Problem.dense(rng.normal(size=n), J, 0.3), ("dense", J)
    edges = [(i, (i + 2) % n, float(rng.normal())) for i in range(n)]
    yield Problem.sparse(rng.normal(size=n), edges, -0.2), ("sparse", edges)
    terms = [((i,), float(rng.normal())) for i in range(n)]
    terms += [(tuple(int(x) for x in rng.choice(n, 3, replace=False)), float(rng.normal())) for _ in range(9)]
    terms += [((1, 1, 4), 0.5), ((2, 2), 0.25)]
    yield Problem.hubo(n, terms, 0.1), ("hubo", terms)


def _energy_by_definition(p, meta, s):
    kind, data = meta
    s = s.astype(float)
    if kind == "dense":
        return p.c + p.h @ s + sum(data[i, j] * s[i] * s[j] for i in range(p.n) for j in range(i + 1, p.n))
    if kind == "sparse":
        return p.c + p.h @ s + sum(v * s[i] * s[j] for i, j, v in data)
    return p.c + sum(c * np.prod([s[v] for v in vs]) for vs, c in data)


def test_energy_delta_fields_by_definition(orc):
    rng = np.random.default_rng(1)
    for p, meta in _rand_problems(rng):
        for bits in itertools.product([-1, 1], repeat=p.n):
            s = np.array(bits, np.int8)
            e = orc.energy(p, s)[0]
            assert abs(e - _energy_by_definition(p, meta, s)) < 1e-9
        s = rng.choice([-1, 1], p.n).astype(np.int8)
        f = orc.local_fields(p, s)
        e0 = orc.energy(p, s)[0]
        for i in range(p.n):
            s2 = s.copy()
            s2[i] = -s2[i]
            d = orc.delta_energy(p, s, i)
            assert abs(d - (orc.energy(p, s2)[0] - e0)) < 1e-9       # dE = E(flip) - E
            assert abs(d - (-2.0 * s[i] * f[i])) < 1e-9              # dE = -2 s f


def test_incremental_fields_walk(orc):
    # 2000-flip incremental-vs-recompute walk (tests/test_higher_order_ising.cpp:146-163 pattern)
    rng = np.random.default_rng(0xC0FFEE)
    for p, _ in _rand_problems(rng):
        s = rng.choice([-1, 1], p.n).astype(np.int8)
        f = orc.local_fields(p, s)
        for _ in range(300):
            i = int(rng.integers(p.n))
            old = int(s[i])
            s[i] = -old
            f = orc.update_fields(p, f, s, i, old)
        assert np.max(np.abs(f - orc.local_fields(p, s))) < 1e-7


def test_qubo_ising_equivalence(orc):
    rng = np.random.default_rng(3)
    n = 6
    Q = rng.integers(-9, 10, (n, n)).astype(float)
    h, J, c = orc.qubo_to_ising(Q)
    p = Problem.dense(h, J, c)
    for bits in itertools.product([0, 1], repeat=n):
        x = np.array(bits, float)
        s = (2 * x - 1).astype(np.int8)  # x = (1 + s) / 2
        assert abs(x @ Q @ x - orc.energy(p, s)[0]) < 1e-9


def _exact_boltzmann_mean(p, orc, beta):
    states = np.array(list(itertools.product([-1, 1], repeat=p.n)), np.int8)
    e = orc.energy(p, states)
    w = np.exp(-beta * (e - e.min()))
    return float((w * e).sum() / w.sum())


def test_sa_stationary_mean(orc):
    # At fixed beta the Metropolis chain samples exp(-beta E); the end-of-step energies are
    # (correlated) samples. A sign error in dE or in the acceptance test moves the mean far away.
    rng = np.random.default_rng(7)
    n = 5
    J = rng.normal(size=(n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    p = Problem.dense(rng.normal(size=n) * 0.5, J)
    beta = 0.7
    tr = np.concatenate([orc.sa_run(p, np.full(4000, beta), 1, seed)["energy_trace"][200:] for seed in range(5)])
    exact = _exact_boltzmann_mean(p, orc, beta)
    assert abs(tr.mean() - exact) < 0.05 * (1 + abs(exact))


def test_sqa_stationary_mean(orc):
    # At fixed (beta, Gamma) the slice-sweep chain samples the Trotter weight
    # exp(-(beta/P) sum_t E(s_t) + J_perp sum_{t,i} s_t,i s_t+1,i) over 2^(N*P) configurations,
    # J_perp = 0.5 ln coth(beta Gamma / P). The engine's trace is the mean slice energy.
    rng = np.random.default_rng(11)
    n, P = 3, 4
    J = rng.normal(size=(n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    p = Problem.dense(rng.normal(size=n) * 0.3, J)
    beta, gamma = 1.5, 1.0
    jp = 0.5 * np.log(1.0 / np.tanh(beta * gamma / P))
    single = np.array(list(itertools.product([-1, 1], repeat=n)), np.int8)
    es = orc.energy(p, single)
    idx = np.array(list(itertools.product(range(2 ** n), repeat=P)))
    E = es[idx]
    spins = single[idx].astype(float)  # [cfg, P, n]
    B = (spins * np.roll(spins, -1, axis=1)).sum(axis=(1, 2))
    logw = -(beta / P) * E.sum(axis=1) + jp * B
    w = np.exp(logw - logw.max())
    exact = float((w * E.mean(axis=1)).sum() / w.sum())
    steps = 6000
    res = [orc.sqa_run(p, np.full(steps, beta), np.full(steps, gamma), P, 4, 1, 0, 0, 0, s)["energy_trace"][300:]
           for s in range(3)]
    est = float(np.concatenate(res).mean())
    assert abs(est - exact) < 0.03 * (1 + abs(exact))
    # a run with the Trotter term's sign flipped would sample a visibly different mean
    logw_bad = -(beta / P) * E.sum(axis=1) - jp * B
    wb = np.exp(logw_bad - logw_bad.max())
    assert abs(float((wb * E.mean(axis=1)).sum() / wb.sum()) - exact) > 0.1


def test_best_energy_matches_best_state(orc):
    # energy(best_state) == best_energy (tests/test_sqa_basic.cpp:30, test_sqa_parallel_tempering.cpp:31)
    for name in ("sa.json", "sqa.json", "sqapt.json", "sqa_chi.json"):
        for c in load(name)["cases"]:
            p = problem(c["problem"])
            o = out(c["out"])
            assert abs(Oracle(ORACLE_SO).energy(p, o["best_state"])[0] - o["best_energy"][0]) < 1e-12


# ---------------------------------------------------------------------------------------------
# 3. randomized comparison with the reference build (only where it exists)
# ---------------------------------------------------------------------------------------------
@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build not present (oracle/_ref)")
def test_random_against_reference(orc):
    ref = Oracle(REF_SO)
    ref.set_threads(1)
    rng = np.random.default_rng(99)
    for case in range(12):
        n = int(rng.integers(3, 24))
        J = rng.normal(size=(n, n))
        J = np.triu(J, 1) + np.triu(J, 1).T
        p = Problem.dense(rng.normal(size=n) * 0.3, J, 0.5)
        b, g = np.linspace(0.3, 3, 10), np.geomspace(3, 0.02, 10)
        for a, r in [(orc.sa_run(p, b, 2, case), ref.sa_run(p, b, 2, case)),
                     (orc.sqa_run(p, b, g, 1 + case % 5, 1 + case % 3, 1, case % 2, (case // 2) % 2, 0, case),
                      ref.sqa_run(p, b, g, 1 + case % 5, 1 + case % 3, 1, case % 2, (case // 2) % 2, 0, case)),
                     (orc.sqapt_run(p, b[:3], g[:3], 2 + case % 3, 1, 1, 4, 1, case % 2, 0, case),
                      ref.sqapt_run(p, b[:3], g[:3], 2 + case % 3, 1, 1, 4, 1, case % 2, 0, case)),
                     (orc.sqa_chi_run(p, 2 + case % 4, 2, 2.0, 3.0, 0.05, 8, 2, 5, 3, 1, seed=case),
                      ref.sqa_chi_run(p, 2 + case % 4, 2, 2.0, 3.0, 0.05, 8, 2, 5, 3, 1, seed=case)),
                     (orc.replica_run(p, b, 2, 1 + case % 4, case), ref.replica_run(p, b, 2, 1 + case % 4, case)),
                     (orc.pt_run(p, b[:2 + case % 4], 1, 6, 1 + case % 2, case),
                      ref.pt_run(p, b[:2 + case % 4], 1, 6, 1 + case % 2, case)),
                     (orc.sqa_opt_run(p, 3, 1 + case % 2, 2.0, 0.05, 2.0, 0.0 if case % 3 else 0.05, 15 / 14, 8, 1,
                                      case % 2, (case // 2) % 2, 4, 2, 0.25, 0.1, case),
                      ref.sqa_opt_run(p, 3, 1 + case % 2, 2.0, 0.05, 2.0, 0.0 if case % 3 else 0.05, 15 / 14, 8, 1,
                                      case % 2, (case // 2) % 2, 4, 2, 0.25, 0.1, case)),
                     (orc.sqapt_opt_run(p, b[:3], np.geomspace(2.0, 0.5, 3), 3, 6, 1, case % 2,
                                        0.0 if case % 3 else 0.05, 15 / 14, 2.0, (case // 2) % 2, 1, 0, 3, 2, case),
                      ref.sqapt_opt_run(p, b[:3], np.geomspace(2.0, 0.5, 3), 3, 6, 1, case % 2,
                                        0.0 if case % 3 else 0.05, 15 / 14, 2.0, (case // 2) % 2, 1, 0, 3, 2, case))]:
            for k in r:
                assert bits_equal(a[k], r[k]), (case, k)
