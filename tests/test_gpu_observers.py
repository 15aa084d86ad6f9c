"""Observers on device runs (annealer.cpp:33-75, sqa_annealer.cpp:87-477, metrics_observer.hpp).

FP64 validation: MetricsObserver / StateTraceObserver records of Annealer::run, SQAMetricsObserver
records of SQAAnnealer::run and the SQAStateTraceObserver tracing path bit-equal to the golden
fixtures written from the reference build. The tracing path runs in FP64 whatever the backend
precision, so the "hip" backend must match the same fixtures. FP32 per-step observers: records
agree with the run's own result (energy trace, magnetization of the recorded states).
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


def test_sa_observers_validation_golden():
    for c in load("observers.json")["sa"]:
        print("sa observer case", c["seed"], flush=True)
        _, m = _model(c["problem"])
        exp = out(c["out"])
        sched = qa.AnnealSchedule.from_betas(dec(c["betas"]).tolist())
        a = qa.Annealer(m, sched, backend="hip_fp64")
        a.set_seed(c["seed"])
        mo = qa.MetricsObserver()
        r = a.run(c["sweeps_per_beta"], mo)
        assert bits_equal(np.array(r.energy_trace), exp["energy_trace"])
        assert bits_equal(np.array(mo.energy_trace), exp["energy_trace"])
        assert bits_equal(np.array(mo.magnetization_trace), exp["magnetization_trace"])
        assert r.best_energy == exp["best_energy"][0] and list(r.best_state.spins) == exp["best_state"].tolist()
        a.set_seed(c["seed"])
        so = qa.StateTraceObserver()
        r2 = a.run(c["sweeps_per_beta"], so)
        assert bits_equal(np.array(r2.energy_trace), exp["energy_trace"])
        assert bits_equal(np.array(so.energy_trace), exp["sweep_energies"])
        assert np.array_equal(np.array(so.state_trace, np.int8).ravel(), exp["sweep_states"].ravel())
        spb, steps = c["sweeps_per_beta"], len(dec(c["betas"]))
        assert list(so.step_trace) == [k // spb for k in range(steps * spb)]
        assert list(so.sweep_trace) == [k % spb for k in range(steps * spb)]


def test_sqa_observers_validation_golden():
    for c in load("observers.json")["sqa"]:
        print("sqa observer case", c["seed"], c["tracing"], flush=True)
        _, m = _model(c["problem"])
        exp = out(c["out"])
        sched = qa.SQASchedule.from_vectors(dec(c["betas"]).tolist(), dec(c["gammas"]).tolist())
        backends = ("hip_fp64", "hip") if c["tracing"] else ("hip_fp64",)
        for backend in backends:
            a = qa.SQAAnnealer(m, sched, c["trotter_slices"], c["replicas"], backend=backend)
            a.set_seed(c["seed"])
            obs = qa.SQAStateTraceObserver() if c["tracing"] else qa.SQAMetricsObserver()
            r = a.run(c["sweeps_per_beta"], c["worldline_sweeps"], c["cluster_sweeps"], 0, obs)
            assert bits_equal(np.array(r.energy_trace), exp["energy_trace"]), backend
            assert r.best_energy == exp["best_energy"][0] and list(r.best_state.spins) == exp["best_state"].tolist()
            if not c["tracing"]:
                assert bits_equal(np.array(obs.energy_trace), exp["energy_trace"])
                assert bits_equal(np.array(obs.magnetization_trace), exp["magnetization_trace"])
                continue
            meta = exp["sweep_meta"].reshape(-1, 4).astype(np.int64)
            assert list(obs.step_trace) == meta[:, 0].tolist() and list(obs.replica_trace) == meta[:, 1].tolist()
            assert list(obs.sweep_trace) == meta[:, 2].tolist() and list(obs.phase_trace) == meta[:, 3].tolist()
            assert bits_equal(np.array(obs.avg_energy_trace), exp["sweep_avg"])
            assert bits_equal(np.array(obs.replica_energy_trace).ravel(), exp["sweep_rep"])
            assert np.array_equal(np.array(obs.state_trace, np.int8).ravel(), exp["sweep_states"].ravel())


def test_observers_fp32_consistent():
    rng = np.random.default_rng(51)
    n = 40
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    m = qa.DenseIsing(np.zeros(n), J)
    a = qa.Annealer(m, qa.AnnealSchedule.linear(0.1, 3.0, 12), backend="hip")
    a.set_seed(2)
    mo = qa.MetricsObserver()
    r = a.run(3, mo)
    assert bits_equal(np.array(mo.energy_trace), np.array(r.energy_trace))
    assert np.all(np.abs(mo.magnetization_trace) <= 1.0)
    assert abs(m.energy(list(r.best_state.spins)) - r.best_energy) < 1e-9
    so = qa.StateTraceObserver()
    so.stride = 2
    a.set_seed(2)
    r2 = a.run(4, so)
    assert list(so.sweep_trace) == [0, 2] * 12
    for st, e in zip(so.state_trace, so.energy_trace):
        assert abs(m.energy(list(st)) - e) < 1e-9  # FP32 chained launches carry exact energies
    s = qa.SQAAnnealer(m, qa.SQASchedule.from_vectors([2.0] * 10, np.geomspace(2.0, 0.05, 10).tolist()), 8, 3,
                       backend="hip")
    s.set_seed(4)
    so2 = qa.SQAMetricsObserver()
    rs = s.run(2, 1, 0, 0, so2)
    assert bits_equal(np.array(so2.energy_trace), np.array(rs.energy_trace))
    assert len(so2.magnetization_trace) == 10 and np.all(np.abs(so2.magnetization_trace) <= 1.0)
    assert abs(m.energy(list(rs.best_state.spins)) - rs.best_energy) < 1e-9
