"""Classical replica annealing and parallel tempering through the drop-in classes on the GPU.

FP64 validation ("hip_fp64"): every field of ReplicaAnnealer::run and
ParallelTemperingAnnealer::run bit-equal to the golden fixtures written from the reference build
(replica_annealer.cpp:46-142, parallel_tempering.cpp:36-122). FP32 ("hip"): returned energies
equal the FP64 energy of the returned states, the averages and the global best follow from the
per-replica results, equal betas accept every swap, and best-energy distributions agree with
the oracle's (two-sample KS, p > 0.01, energies rounded to 1e-9 so that exact ties stay ties).
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
    # the Python ctor takes no constant: a degree-0 term folds into it (higher_order_ising.cpp:10-90)
    return p, qa.HigherOrderIsing([(list(v), c) for v, c in p.terms] + [([], p.c)], p.n)


def _sk(rng, n):
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def test_replica_validation_golden():
    for c in load("replica.json")["cases"]:
        print("replica case seed", c["seed"], flush=True)
        _, m = _model(c["problem"])
        a = qa.ReplicaAnnealer(m, qa.AnnealSchedule.from_betas(dec(c["betas"]).tolist()), c["replicas"],
                               backend="hip_fp64")
        a.set_seed(c["seed"])
        r = a.run(c["sweeps_per_beta"])
        exp = out(c["out"])
        R = c["replicas"]
        got = dict(best_states=np.array([list(x.best_state.spins) for x in r.replicas], np.int8),
                   best_energies=np.array([x.best_energy for x in r.replicas]),
                   energy_traces=np.concatenate([x.energy_trace for x in r.replicas]),
                   magnetization_traces=np.concatenate([x.magnetization_trace for x in r.replicas]),
                   global_best_state=np.array(r.global_best_state.spins, np.int8),
                   global_best_energy=np.array([r.global_best_energy]),
                   average_energy_trace=np.array(r.average_energy_trace),
                   average_magnetization_trace=np.array(r.average_magnetization_trace))
        assert len(r.replicas) == R
        for k in exp:
            assert bits_equal(np.ravel(got[k]), np.ravel(exp[k])), (c["seed"], k)


def test_pt_validation_golden():
    for c in load("pt.json")["cases"]:
        print("pt case seed", c["seed"], flush=True)
        _, m = _model(c["problem"])
        a = qa.ParallelTemperingAnnealer(m, dec(c["betas"]).tolist(), backend="hip_fp64")
        a.set_seed(c["seed"])
        r = a.run(c["sweeps_per_step"], c["steps"], c["swap_interval"])
        exp = out(c["out"])
        got = dict(final_states=np.array([list(s.spins) for s in r.final_states], np.int8),
                   final_energies=np.array(r.final_energies), best_state=np.array(r.best_state.spins, np.int8),
                   best_energy=np.array([r.best_energy]), average_energy_trace=np.array(r.average_energy_trace),
                   swap_acceptance_trace=np.array(r.swap_acceptance_trace))
        for k in exp:
            assert bits_equal(np.ravel(got[k]), np.ravel(exp[k])), (c["seed"], k)


def test_pt_validation_consecutive_runs_continue_engine():
    """Two runs without reseeding continue the annealer's engine exactly as the reference's do."""
    from oracle.orc import Oracle, ORACLE_SO
    c = load("pt.json")["cases"][0]
    p, m = _model(c["problem"])
    betas = dec(c["betas"]).tolist()
    a = qa.ParallelTemperingAnnealer(m, betas, backend="hip_fp64")
    a.set_seed(c["seed"])
    first = a.run(c["sweeps_per_step"], c["steps"], c["swap_interval"])
    # run 1 equals the fixture; only the fact that run 2 differs from a fresh seeded run and is
    # itself deterministic is checked here (the oracle has no two-run entry point)
    b = qa.ParallelTemperingAnnealer(m, betas, backend="hip_fp64")
    b.set_seed(c["seed"])
    assert bits_equal(np.array(b.run(c["sweeps_per_step"], c["steps"], c["swap_interval"]).average_energy_trace),
                      np.array(first.average_energy_trace))
    second_a = a.run(c["sweeps_per_step"], c["steps"], c["swap_interval"])
    second_b = b.run(c["sweeps_per_step"], c["steps"], c["swap_interval"])
    assert bits_equal(np.array(second_a.average_energy_trace), np.array(second_b.average_energy_trace))
    assert not bits_equal(np.array(second_a.average_energy_trace), np.array(first.average_energy_trace))
    o = Oracle(ORACLE_SO).pt_run(p, betas, c["sweeps_per_step"], c["steps"], c["swap_interval"], c["seed"])
    assert bits_equal(np.array(first.final_energies), o["final_energies"])


def test_replica_fp32_invariants_and_distribution():
    from scipy.stats import ks_2samp
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(31)
    n, R = 40, 256
    J, h = _sk(rng, n), rng.normal(0, 0.1, n)
    m = qa.DenseIsing(h, J)
    betas = np.linspace(0.1, 3.0, 24)
    a = qa.ReplicaAnnealer(m, qa.AnnealSchedule.from_betas(betas.tolist()), R, backend="hip")
    a.set_seed(4)
    r = a.run(2)
    best = np.array([x.best_energy for x in r.replicas])
    for x in r.replicas[:32]:
        assert abs(m.energy(list(x.best_state.spins)) - x.best_energy) < 1e-9
        assert np.all(np.abs(x.magnetization_trace) <= 1.0)
        assert np.allclose(np.array(x.magnetization_trace) * n, np.round(np.array(x.magnetization_trace) * n))
    assert r.global_best_energy == best.min()
    assert abs(m.energy(list(r.global_best_state.spins)) - r.global_best_energy) < 1e-9
    et = np.array([x.energy_trace for x in r.replicas])
    acc = np.zeros(len(betas))
    for row in et:
        acc = acc + row
    assert bits_equal(acc / R, np.array(r.average_energy_trace))
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    cpu = orc.replica_run(Problem.dense(h, J), betas, 2, R, 77)["best_energies"]
    p = ks_2samp(np.round(best, 9), np.round(cpu, 9)).pvalue
    print("replica KS p", p, "gpu mean", best.mean(), "cpu mean", cpu.mean(), flush=True)
    assert p > 0.01


def test_pt_fp32_invariants_and_distribution():
    from scipy.stats import ks_2samp
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(32)
    n = 32
    J = _sk(rng, n)
    m = qa.DenseIsing(np.zeros(n), J)
    betas = np.geomspace(0.3, 3.0, 8).tolist()
    a = qa.ParallelTemperingAnnealer(m, betas, backend="hip")
    a.set_seed(9)
    r = a.run(2, 20, 2)
    fe = np.array(r.final_energies)
    for s, e in zip(r.final_states, fe):
        assert abs(m.energy(list(s.spins)) - e) < 1e-9
    assert abs(m.energy(list(r.best_state.spins)) - r.best_energy) < 1e-9
    assert r.best_energy <= fe.min()
    sw = np.array(r.swap_acceptance_trace)
    assert np.all(sw[0::2] == 0.0) and np.all((sw >= 0) & (sw <= 1))
    # equal betas: every swap has delta 0 and is accepted (parallel_tempering.cpp:99)
    eq = qa.ParallelTemperingAnnealer(m, [1.1] * 5, backend="hip")
    eq.set_seed(2)
    assert np.all(np.array(eq.run(1, 6, 1).swap_acceptance_trace) == 1.0)
    # best energies over independent seeds against the oracle's
    gpu, cpu = [], []
    orc = Oracle(ORACLE_SO)
    orc.set_threads(1)
    p = Problem.dense(np.zeros(n), J)
    for k in range(96):
        a.set_seed(1000 + k)
        gpu.append(a.run(1, 12, 1).best_energy)
        cpu.append(orc.pt_run(p, betas, 1, 12, 1, 5000 + k)["best_energy"])
    pv = ks_2samp(np.round(gpu, 9), np.round(cpu, 9)).pvalue
    print("pt KS p", pv, "gpu mean", np.mean(gpu), "cpu mean", np.mean(cpu), flush=True)
    assert pv > 0.01
