"""The drop-in API on the GPU: the reference's own C++ test assertions through the pybind11
classes, solve() on every in-scope method, and FP64 validation runs through the classes equal to
the oracle's runs of the same seed (the classes draw the reference's random stream on the host)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

qa = pytest.importorskip("qlatanneal_amd")
from oracle.orc import Oracle, Problem, ORACLE_SO  # noqa: E402


def _sk(rng, n):
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def test_reference_sqa_basic():
    # tests/test_sqa_basic.cpp:8-31
    d = qa.DenseIsing(np.zeros(3), np.array([[0, .5, -.3], [.5, 0, .2], [-.3, .2, 0]]))
    for backend in ("hip", "hip_fp64"):
        a = qa.SQAAnnealer(d, qa.SQASchedule.from_vectors([0.2, 0.5, 1.0], [2.0, 1.0, 0.5]), 8, 2, backend=backend)
        a.set_seed(42)
        r = a.run(5, 1)
        assert len(r.energy_trace) == 3
        assert abs(d.energy(r.best_state.spins) - r.best_energy) < 1e-12


def test_reference_sqa_parallel_tempering():
    # tests/test_sqa_parallel_tempering.cpp:7-32
    d = qa.DenseIsing(np.zeros(3), np.array([[0, .5, -.2], [.5, 0, .3], [-.2, .3, 0]]))
    for backend in ("hip", "hip_fp64"):
        a = qa.SQAParallelTemperingAnnealer(d, [0.2, 0.5, 1.0], [2.0, 1.0, 0.5], 8, backend=backend)
        a.set_seed(123)
        r = a.run(4, 1, 5, 1, 1, 0)
        assert len(r.final_states) == 3 and len(r.final_energies) == 3
        assert len(r.average_energy_trace) == 5 and len(r.swap_acceptance_trace) == 5
        assert abs(d.energy(r.best_state.spins) - r.best_energy) < 1e-12


def test_reference_chi_exact_ground_states():
    # tests/test_sqa_chi_annealer.cpp:241-271
    n = 12
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = J[i + 1, i] = -1.0
    chain = qa.DenseIsing(np.zeros(n), J)
    ring = qa.SparseIsing(np.zeros(n), [qa.SparseEdge(i, (i + 1) % n, -1.0) for i in range(n)], n)
    for backend in ("hip", "hip_fp64"):
        a = qa.SQAChiAnnealer(chain, 8, 4, backend=backend)
        a.set_seed(7)
        r = a.run_chi(4.0, 3.0, 0.05, 100, 10, 10, 20, 6)
        assert r.best_energy == -(n - 1)
        g = np.array(r.gamma_schedule)
        assert np.all(np.diff(g) <= 1e-12) and abs(g[0] - 3.0) < 1e-9 and abs(g[-1] - 0.05) < 1e-6
        b = qa.SQAChiAnnealer(ring, 8, 4, backend=backend)
        b.set_seed(11)
        assert b.run_chi(4.0, 3.0, 0.05, 100, 10, 10, 20, 6).best_energy == -n


def test_classes_fp64_equal_oracle():
    rng = np.random.default_rng(5)
    n = 20
    J, h = _sk(rng, n), rng.normal(0, 0.2, n)
    d = qa.DenseIsing(h, J, 0.5)
    p = Problem.dense(h, J, 0.5)
    orc = Oracle(ORACLE_SO)
    orc.set_threads(1)
    betas, gammas = np.linspace(0.3, 3, 12), np.geomspace(3, 0.02, 12)
    a = qa.Annealer(d, qa.AnnealSchedule.from_betas(betas.tolist()), backend="hip_fp64")
    a.set_seed(3)
    r = a.run(4)
    o = orc.sa_run(p, betas, 4, 3)
    assert r.best_energy == o["best_energy"] and list(r.best_state.spins) == o["best_state"].tolist()
    assert np.array_equal(r.energy_trace, o["energy_trace"])
    # consecutive runs without reseeding continue the engine as the reference does: the second
    # and third runs equal the oracle's second and third runs on one engine
    s = qa.SQAAnnealer(d, qa.SQASchedule.from_vectors(betas.tolist(), gammas.tolist()), 4, 3, backend="hip_fp64")
    s.set_seed(8)
    r1 = s.run(2, 1, 1)
    r2 = s.run(2, 1, 1)
    r3 = s.run(2, 1, 1)
    o1 = orc.sqa_run(p, betas, gammas, 4, 3, 2, 1, 1, 0, 8)
    assert r1.best_energy == o1["best_energy"] and np.array_equal(r1.energy_trace, o1["energy_trace"])
    for k, r in ((2, r2), (3, r3)):
        o = orc.sqa_run_repeat(p, k, betas, gammas, 4, 3, 2, 1, 1, 0, 8)
        assert r.best_energy == o["best_energy"], k
        assert list(r.best_state.spins) == o["best_state"].tolist(), k
        assert np.array_equal(r.energy_trace, o["energy_trace"]), k
    assert not np.array_equal(r2.energy_trace, r1.energy_trace)
    pt = qa.SQAParallelTemperingAnnealer(d, betas[:4].tolist(), gammas[:4].tolist(), 3, backend="hip_fp64")
    pt.set_seed(4)
    rp = pt.run(2, 1, 6, 1, 1, 0)
    op = orc.sqapt_run(p, betas[:4], gammas[:4], 3, 2, 1, 6, 1, 1, 0, 4)
    assert rp.best_energy == op["best_energy"] and np.array_equal(rp.average_energy_trace, op["average_energy_trace"])
    ch = qa.SQAChiAnnealer(d, 4, 2, backend="hip_fp64")
    ch.set_seed(6)
    rc = ch.run_chi(2.0, 3.0, 0.05, 10, 2, 5, 3, 1)
    oc = orc.sqa_chi_run(p, 4, 2, 2.0, 3.0, 0.05, 10, 2, 5, 3, 1, seed=6)
    assert rc.best_energy == oc["best_energy"] and np.array_equal(rc.scan_chi_B, oc["scan_chi_B"])
    assert np.array_equal(rc.gamma_schedule, oc["gamma_schedule"])


def test_solve_all_methods():
    rng = np.random.default_rng(12)
    n = 16
    Q = rng.integers(-10, 11, (n, n)).astype(float)
    Q = (Q + Q.T) / 2
    ising = qa.QUBO(Q).to_ising()
    for method, kw in [("sa", {}), ("sqa", dict(trotter_slices=8, replicas=2)),
                       ("sqapt", dict(trotter_slices=8, replicas=4, pt_steps=10)),
                       ("sqa_chi", dict(trotter_slices=8, replicas=2, chi_scan_points=6, chi_scan_sweeps=4,
                                        chi_scan_burn=2, sweeps_per_beta=4))]:
        res = qa.solve(Q, method=method, reads=3, seed=1, return_bits=True, progress=False, **kw)
        assert res.method == method and len(res.samples) == 3
        for smp, e in zip(res.samples, res.energies):
            x = np.asarray(smp, float)
            assert set(np.unique(smp)) <= {0, 1}
            assert abs(x @ Q @ x - e) < 1e-9  # returned energy is the exact QUBO energy of the sample
        assert res.best_energy == min(res.energies)
        spins = 2 * np.asarray(res.best_sample) - 1
        assert abs(ising.energy(spins.tolist()) - res.best_energy) < 1e-9


def test_solve_seeded_reads_reproducible():
    rng = np.random.default_rng(13)
    J = _sk(rng, 24)
    d = qa.DenseIsing(np.zeros(24), J)
    a = qa.solve(d, method="sqa", reads=2, seed=5, progress=False, trotter_slices=8)
    b = qa.solve(d, method="sqa", reads=2, seed=5, progress=False, trotter_slices=8)
    assert a.energies == b.energies


def test_solve_sa_reads_batched_equal_sequential():
    """solve(method="sa", seed=s) runs its reads in one launch; in validation mode every read equals
    set_seed(s + r) followed by Annealer.run, as the reference's read loop (solver.py:679) does."""
    rng = np.random.default_rng(15)
    n = 20
    J = _sk(rng, n)
    d = qa.DenseIsing(rng.normal(0, 0.2, n), J)
    sched = qa.auto_schedule_sa()
    res = qa.solve(d, method="sa", reads=5, seed=11, backend="hip_fp64", progress=False, sweeps_per_beta=10,
                   schedule=sched)
    a = qa.Annealer(d, sched, backend="hip_fp64")
    for r in range(5):
        a.set_seed(11 + r)
        one = a.run(10)
        assert res.energies[r] == one.best_energy
        assert list(res.samples[r]) == list(one.best_state.spins)
    fast = qa.solve(d, method="sa", reads=8, seed=3, progress=False)
    assert len(fast.energies) == 8 and fast.best_energy == min(fast.energies)
    for s_, e in zip(fast.samples, fast.energies):
        assert abs(d.energy(np.asarray(s_).astype(int).tolist()) - e) < 1e-9


def test_sa_fp32_reads():
    rng = np.random.default_rng(14)
    n = 64
    J = _sk(rng, n)
    d = qa.DenseIsing(np.zeros(n), J)
    betas = (0.1 * 30.0 ** (np.arange(100) / 99.0)).tolist()
    a = qa.Annealer(d, qa.AnnealSchedule.from_betas(betas), backend="hip")
    res = a.run_reads(10, 32, 1)
    orc = Oracle(ORACLE_SO)
    p = Problem.dense(np.zeros(n), J)
    cpu = [orc.sa_run(p, betas, 10, 100 + k)["best_energy"] for k in range(32)]
    gpu = [r.best_energy for r in res]
    for r in res:
        assert r.best_energy == d.energy(r.best_state.spins)
    assert abs(np.mean(gpu) - np.mean(cpu)) < 0.05 * abs(np.mean(cpu))
