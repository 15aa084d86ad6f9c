"""The reference's own test assertions for the engines built on the device path, run through the
drop-in classes on the GPU in both precisions:

  tests/test_replica_annealer.cpp, tests/test_parallel_tempering.cpp,
  tests/test_sqa_optimal_schedule.cpp, tests/test_sqapt_optimal_schedule.cpp,
  tests/test_optimal_schedule_resolution.py (the run_optimal part).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

qa = pytest.importorskip("qlatanneal_amd")

BACKENDS = ("hip", "hip_fp64")
M64 = (1 << 64) - 1


def _three_spin():
    h = np.array([0.1, -0.1, 0.2])
    J = np.array([[0.0, 0.4, 0.0], [0.4, 0.0, -0.2], [0.0, -0.2, 0.0]])
    return qa.DenseIsing(h, J)


def _xorshift_dense(n):
    """The instance of test_sqa_optimal_schedule.cpp:24-40 (xorshift64 couplings in [-3, 3))."""
    s = 0x9E3779B97F4A7C15
    J = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            s ^= (s << 13) & M64
            s ^= s >> 7
            s ^= (s << 17) & M64
            w = ((s % 100000) / 100000.0 - 0.5) * 6.0
            J[i, j] = J[j, i] = w
    return qa.DenseIsing(np.zeros(n), J)


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_replica_annealer(backend):
    # tests/test_replica_annealer.cpp:7-27
    sched = qa.AnnealSchedule.linear(0.2, 1.0, 4)
    a = qa.ReplicaAnnealer(_three_spin(), sched, 3, backend=backend)
    a.set_seed(123)
    r = a.run(5)
    assert len(r.replicas) == 3
    assert len(r.average_energy_trace) == len(sched.betas)
    assert len(r.average_magnetization_trace) == len(sched.betas)


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_parallel_tempering(backend):
    # tests/test_parallel_tempering.cpp:7-30
    d = _three_spin()
    a = qa.ParallelTemperingAnnealer(d, [0.2, 0.6, 1.0], backend=backend)
    a.set_seed(7)
    r = a.run(3, 4, 1)
    assert len(r.average_energy_trace) == 4 and len(r.swap_acceptance_trace) == 4
    assert abs(d.energy(r.best_state.spins) - r.best_energy) < 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_sqa_optimal_schedule(backend):
    # tests/test_sqa_optimal_schedule.cpp:44-113
    ham = _xorshift_dense(10)
    sched = qa.SQASchedule.from_vectors([1.0], [1.0])
    steps, j0, j1 = 150, 0.5, 5.0
    cal = qa.SQAAnnealer(ham, sched, 8, 4, backend=backend)
    cal.set_seed(2024)
    rc = cal.run_optimal(1.0, j0, j1, 0.0, 15.0 / 14.0, steps, 8, 0, 0, 12, 8)
    assert rc.calibrated_eps_tilde > 0.0
    assert rc.resolved_j_perp_end == j1 and abs(rc.j_perp_start - j0) < 1e-9
    assert len(rc.j_perp_trace) == steps and len(rc.chi_B_trace) == steps
    jt = np.array(rc.j_perp_trace)
    assert abs(jt[0] - rc.j_perp_start) < 1e-9 and np.all(np.diff(jt) >= -1e-12)
    frac = (rc.final_j_perp - rc.j_perp_start) / (j1 - j0)
    assert frac > 0.9
    assert max(rc.chi_B_trace) > min(rc.chi_B_trace)
    fix = qa.SQAAnnealer(ham, sched, 8, 4, backend=backend)
    fix.set_seed(2024)
    rf = fix.run_optimal(1.0, j0, j1, 1e-4, 15.0 / 14.0, steps, 8, 0, 0)
    frac_fixed = (rf.final_j_perp - rf.j_perp_start) / (j1 - j0)
    assert frac_fixed < 0.5 and frac > frac_fixed + 0.3


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_sqapt_optimal_schedule(backend):
    # tests/test_sqapt_optimal_schedule.cpp:47-121
    ham = _xorshift_dense(10)
    betas, gammas = [0.2, 0.5, 1.0, 2.0], [2.0, 1.0, 0.5, 0.25]
    steps, j1 = 150, 4.0
    cal = qa.SQAParallelTemperingAnnealer(ham, betas, gammas, 8, backend=backend)
    cal.set_seed(2024)
    rc = cal.run_optimal(steps, 8, 0, 0.0, 15.0 / 14.0, j1, 0, 1, 0, 12, 8)
    assert rc.calibrated_eps_tilde > 0.0 and rc.resolved_j_perp_end == j1
    assert len(rc.j_perp_trace) == steps and len(rc.chi_B_trace) == steps
    jt = np.array(rc.j_perp_trace)
    assert abs(jt[0] - rc.j_perp_start) < 1e-9 and np.all(np.diff(jt) >= -1e-12)
    rng_ = j1 - rc.j_perp_start
    assert rng_ > 0.5
    frac = (rc.final_j_perp - rc.j_perp_start) / rng_
    assert frac > 0.9
    assert max(rc.chi_B_trace) > min(rc.chi_B_trace)
    fix = qa.SQAParallelTemperingAnnealer(ham, betas, gammas, 8, backend=backend)
    fix.set_seed(2024)
    rf = fix.run_optimal(steps, 8, 0, 1e-4, 15.0 / 14.0, j1, 0, 1, 0)
    frac_fixed = (rf.final_j_perp - rf.j_perp_start) / rng_
    assert frac_fixed < 0.5 and frac > frac_fixed + 0.3


def _resolved_end(J, j_perp_end, backend):
    # tests/test_optimal_schedule_resolution.py:93-105
    n = J.shape[0]
    ann = qa.SQAParallelTemperingAnnealer(qa.DenseIsing(np.zeros(n), J), [0.2, 0.5, 1.0], [2.0, 1.0, 0.5], 16,
                                          backend=backend)
    ann.set_seed(7)
    res = ann.run_optimal(num_steps=2, sweeps_per_step=2, worldline_sweeps=0, eps_tilde=0.0, j_perp_end=j_perp_end,
                          calib_probes=4, calib_sweeps=2)
    return res.resolved_j_perp_end, res.j_perp_start


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_optimal_schedule_resolution(backend):
    # tests/test_optimal_schedule_resolution.py:108-138 (instances of :45-84)
    rng = np.random.default_rng(1)
    n = 40
    W = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < 0.5:
                W[i, j] = W[j, i] = rng.uniform(1.0, 10.0)
    target = 5.0 * qa.j_rms_from_problem(W)
    resolved, start = _resolved_end(W, target, backend)
    assert target > start and resolved == pytest.approx(target, rel=1e-9)
    rng = np.random.default_rng(2)
    S = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            S[i, j] = S[j, i] = float(rng.normal(0.0, 1.0 / math.sqrt(n)))
    _, start = _resolved_end(S, 1.0, backend)
    target = max(5.0 * qa.j_rms_from_problem(S), start + 2.0)
    resolved, _ = _resolved_end(S, target, backend)
    assert resolved == pytest.approx(target, rel=1e-9)
    rng = np.random.default_rng(0)
    M = np.zeros((30, 30))
    for i in range(30):
        for _ in range(3):
            j = int(rng.integers(0, 30))
            if j != i:
                M[i, j] = M[j, i] = 0.5
    resolved, start = _resolved_end(M, 0.0, backend)
    assert resolved > start + 1e5
