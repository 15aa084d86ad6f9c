"""FP32 production mode with more than 64 Trotter slices per replica.

The reference accepts any slice count (sqa_annealer.cpp:28-30) and continuous_time_slices raises it
past trotter_slices (sqa_annealer.cpp:97-99). Past 64 slices the two-barrier kernel runs one
replica per workgroup with one thread per slice over several wavefronts (multi-word flip masks,
multi-word cluster masks). Checks: every returned energy is the oracle's FP64 energy of the
returned state (dense, sparse, HUBO; slice, worldline and cluster moves; SQA, SQA-PT, sqa_chi);
the LDS-state and global-state kernels agree bit for bit (the global path's one-step-ahead
prefetch takes the previous slice's spin from memory at a wavefront boundary); and best-energy
distributions of independent reads agree with the oracle SQA (two-sample KS, p > 0.01).
"""
import numpy as np
import pytest

from tests.golden_io import bits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from qlatanneal_amd.abi import Context
    return Context(0)


def _sk(rng, n):
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def test_many_slices_exact_energies(ctx):
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    from tests.golden_io import consolidate_hubo
    from tests.test_gpu_parity import _sparse_case, _hubo_case
    orc = Oracle(ORACLE_SO)
    rng = np.random.default_rng(65)
    n = 40
    h, J = rng.normal(0, 0.1, n), _sk(rng, n)
    hs, edges = _sparse_case(rng, 90, 0.05)
    terms = _hubo_case(rng, 30, 80)
    hh, off, var, co, dc = consolidate_hubo(30, terms)
    cases = (("dense", Problem.dense(h, J, 0.5), abi.DeviceProblem.dense(ctx, h, J, 0.5)),
             ("sparse", Problem.sparse(hs, edges, 0.5), abi.DeviceProblem.sparse(ctx, hs, edges, 0.5)),
             ("hubo", Problem.hubo(30, terms, -0.25), abi.DeviceProblem.hubo(ctx, 30, hh, off, var, co, -0.25 + dc)))
    for name, po, pd in cases:
        for P, R, wl, cl in ((65, 2, 0, 0), (96, 3, 1, 1), (130, 1, 1, 1), (384, 1, 0, 1)):
            print("many-slices exact-energy case", name, P, R, wl, cl, flush=True)
            r = abi.sqa_run(pd, np.full(10, 3.0), np.geomspace(2.0, 0.05, 10), P, R, 1, wl, cl, abi.MODE_FP32, 13)
            assert r["best_energy"] == orc.energy(po, r["best_state"])[0], (name, P, R, wl, cl)
            assert np.all(np.isfinite(r["energy_trace"])) and 0.0 <= r["accept_rate"] <= 1.0


def test_many_slices_state_paths_bit_identical(ctx, monkeypatch):
    """LDS state vs global state (float4 and column-loop updates), slice sweeps alone and with
    worldline and cluster moves."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(66)
    for n, P, R, wl, cl in [(64, 96, 2, 0, 0), (100, 130, 1, 0, 0), (48, 384, 1, 0, 0), (64, 128, 2, 1, 1),
                            (36, 200, 1, 1, 1)]:
        print("many-slices state case", n, P, R, wl, cl, flush=True)
        dp = abi.DeviceProblem.dense(ctx, rng.normal(0, 0.2, n), _sk(rng, n), 0.25)
        betas, gammas = np.geomspace(0.3, 3.0, 4), np.geomspace(3.0, 0.05, 4)
        res = []
        for state, f4, sl in (("lds", "1", "1"), ("global", "1", "1"), ("global", "1", "0"), ("global", "0", "1")):
            monkeypatch.setenv("QANNEAL_SQA_STATE", state)
            monkeypatch.setenv("QANNEAL_SQA_GLOBAL_F4", f4)
            monkeypatch.setenv("QANNEAL_SQA_SPINS_LDS", sl)
            res.append(abi.sqa_run(dp, betas, gammas, P, R, 2, wl, cl, abi.MODE_FP32, 31))
        for r in res[1:]:
            assert r["accept_rate"] == res[0]["accept_rate"], (n, P, R)
            assert np.array_equal(r["best_state"], res[0]["best_state"]), (n, P, R)
            assert bits_equal([r["best_energy"]], [res[0]["best_energy"]]), (n, P, R)
            assert bits_equal(r["energy_trace"], res[0]["energy_trace"]), (n, P, R)
    monkeypatch.delenv("QANNEAL_SQA_STATE", raising=False)
    monkeypatch.delenv("QANNEAL_SQA_GLOBAL_F4", raising=False)
    monkeypatch.delenv("QANNEAL_SQA_SPINS_LDS", raising=False)


def test_many_slices_distribution_vs_oracle(ctx):
    from scipy.stats import ks_2samp
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(67)
    n, P, steps = 16, 96, 24
    J = _sk(rng, n)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    betas, gammas = np.full(steps, 2.0), np.linspace(2.0, 0.05, steps)
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    for wl, cl in ((0, 0), (1, 1)):
        gpu = np.array([abi.sqa_run(dp, betas, gammas, P, 1, 1, wl, cl, abi.MODE_FP32, 2000 + k)["best_energy"]
                        for k in range(128)])
        cpu = np.array([orc.sqa_run(Problem.dense(np.zeros(n), J), betas, gammas, P, 1, 1, wl, cl, 0,
                                    7000 + k)["best_energy"] for k in range(128)])
        p = ks_2samp(gpu, cpu).pvalue
        print("KS many slices", wl, cl, p, "gpu mean", gpu.mean(), "cpu mean", cpu.mean(), flush=True)
        assert p > 0.01, (wl, cl)


def test_many_slices_sqapt_chi_and_continuous_time(ctx, monkeypatch):
    import qlatanneal_amd as qa
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    orc = Oracle(ORACLE_SO)
    rng = np.random.default_rng(68)
    n = 48
    h, J = rng.normal(0, 0.1, n), _sk(rng, n)
    po, dp = Problem.dense(h, J, 0.25), abi.DeviceProblem.dense(ctx, h, J, 0.25)
    print("many-slices SQA-PT", flush=True)
    r = abi.sqapt_run(dp, np.geomspace(0.5, 4.0, 5), np.geomspace(3.0, 0.05, 5), 80, 2, 1, 5, 1, 1,
                      ladders=2, mode=abi.MODE_FP32, seed=3)
    for lad in range(2):
        assert r["best_energy"][lad] == orc.energy(po, r["best_state"][lad])[0], lad
        for k, e in enumerate(r["final_energies"][lad]):
            assert e == orc.energy(po, r["final_states"][lad][k])[0], (lad, k)
    print("many-slices sqa_chi", flush=True)
    kw = dict(beta=2.5, gamma_start=3.0, gamma_end=0.05, num_steps=5, sweeps_per_step=2, scan_points=4,
              scan_sweeps=2, scan_burn=1, seed=37)
    res = []
    for v in ("1", "0"):
        monkeypatch.setenv("QANNEAL_CHI_WAVE", v)
        res.append(abi.sqa_chi_run(dp, 97, 2, **kw))
    monkeypatch.delenv("QANNEAL_CHI_WAVE", raising=False)
    a, b = res
    assert a["best_energy"] == orc.energy(po, a["best_state"])[0]
    assert np.array_equal(a["best_state"], b["best_state"]) and bits_equal(a["energy_trace"], b["energy_trace"])
    assert bits_equal(a["scan_chi_B"], b["scan_chi_B"])
    print("continuous_time_slices through SQAAnnealer", flush=True)
    ham = qa.DenseIsing(h, J, 0.25)
    sched = qa.SQASchedule.from_vectors(np.full(8, 3.0).tolist(), np.geomspace(2.0, 0.05, 8).tolist())
    ann = qa.SQAAnnealer(ham, sched, 8, 2, backend="hip")
    ann.set_seed(5)
    res = ann.run(1, 0, 1, 100)
    assert res.best_energy == ham.energy(res.best_state.spins)
    assert len(res.energy_trace) == 8


def test_slices_beyond_kernel_limit_refused(ctx):
    from qlatanneal_amd import abi
    rng = np.random.default_rng(69)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(8), _sk(rng, 8))
    with pytest.raises(abi.QaError, match="384 Trotter slices"):
        abi.sqa_run(dp, np.full(2, 1.0), np.full(2, 1.0), 385, 1, 1, 0, 0, abi.MODE_FP32, 1)
