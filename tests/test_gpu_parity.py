"""Device parity through the C-ABI (include/qanneal_hip.h).

FP64 validation mode: bit-exact against the golden fixtures written from the reference build
(best_state, best_energy and every energy_trace entry). FP32 production mode: the energy returned
with each sample equals the FP64 recompute exactly, and residual-energy distributions agree with
the oracle CPU path (two-sample KS, p > 0.01).
"""
import numpy as np
import pytest

from tests.golden_io import load, out, dec, bits_equal, device_problem, problem
from tests.mt64 import MT64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from qlatanneal_amd.abi import Context
    return Context(0)


def test_energy_kernel_golden(ctx):
    for c in load("hamiltonian.json")["cases"]:
        dp = device_problem(ctx, c["problem"])
        e = dp.energy(np.array(c["spins"], np.int8))
        assert bits_equal(e, dec(c["energy"]))


def test_sa_validation_golden(ctx):
    from qlatanneal_amd import abi
    for c in load("sa.json")["cases"]:
        print("sa case seed", c["seed"], flush=True)
        dp = device_problem(ctx, c["problem"])
        betas = dec(c["betas"])
        g = MT64(c["seed"])
        init = g.spins(dp.n)
        uni = g.uniforms(len(betas) * c["sweeps_per_beta"] * dp.n)
        r = abi.sa_run(dp, betas, c["sweeps_per_beta"], 1, abi.MODE_FP64_VALIDATION, 0, init, uni)
        exp = out(c["out"])
        assert bits_equal(r["best_states"][0], exp["best_state"])
        assert bits_equal(r["best_energies"][:1], exp["best_energy"])
        assert bits_equal(r["energy_traces"][0], exp["energy_trace"])


def test_sqa_validation_golden(ctx):
    from qlatanneal_amd import abi
    for c in load("sqa.json")["cases"]:
        print("sqa case", c["seed"], c["trotter_slices"], c["replicas"], flush=True)
        dp = device_problem(ctx, c["problem"])
        P = max(c["trotter_slices"], c["continuous_time_slices"])
        R = c["replicas"]
        g = MT64(c["seed"])
        init = g.spins(R * P * dp.n)
        seeds = g.raw(R)
        r = abi.sqa_run(dp, dec(c["betas"]), dec(c["gammas"]), P, R, c["sweeps_per_beta"],
                        c["worldline_sweeps"], c["cluster_sweeps"], abi.MODE_FP64_VALIDATION, 0, init, seeds)
        exp = out(c["out"])
        assert bits_equal(r["best_state"], exp["best_state"]), c["seed"]
        assert bits_equal([r["best_energy"]], exp["best_energy"]), c["seed"]
        assert bits_equal(r["energy_trace"], exp["energy_trace"]), c["seed"]


def _sk(rng, n):
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def test_sqa_fp32_returned_energy_exact(ctx):
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(3)
    orc = Oracle(ORACLE_SO)
    for n, P, R, wl, cl in [(40, 8, 4, 0, 0), (33, 5, 3, 1, 1), (20, 1, 2, 0, 0), (64, 32, 2, 1, 0)]:
        print("fp32 exact-energy case", n, P, R, flush=True)
        J = _sk(rng, n)
        h = rng.normal(0, 0.1, n)
        dp = abi.DeviceProblem.dense(ctx, h, J, 0.5)
        r = abi.sqa_run(dp, np.full(30, 4.0), np.geomspace(3, 0.01, 30), P, R, 1, wl, cl, abi.MODE_FP32, 11)
        e_host = orc.energy(Problem.dense(h, J, 0.5), r["best_state"])[0]
        assert r["best_energy"] == e_host
        assert np.all(np.isfinite(r["energy_trace"]))


def test_sqa_fp32_distribution_vs_oracle(ctx):
    from scipy.stats import ks_2samp
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(17)
    n, P, steps = 48, 8, 60
    J = _sk(rng, n)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    betas, gammas = np.full(steps, 2.0), np.linspace(2.0, 0.01, steps)
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    # 256 independent reads per path (R=1 per read, as solve() runs them)
    gpu = np.array([abi.sqa_run(dp, betas, gammas, P, 1, 1, 0, 0, abi.MODE_FP32, 1000 + k)["best_energy"]
                    for k in range(256)])
    cpu = np.array([orc.sqa_run(Problem.dense(np.zeros(n), J), betas, gammas, P, 1, 1, 0, 0, 0, 5000 + k)["best_energy"]
                    for k in range(256)])
    p = ks_2samp(gpu, cpu).pvalue
    print("KS p", p, "gpu mean", gpu.mean(), "cpu mean", cpu.mean(), flush=True)
    assert p > 0.01


def test_sqa_fp32_stationary_mean(ctx):
    """At fixed (beta, Gamma) the FP32 sweep must sample the Trotter weight
    exp(-(beta/P) sum_t E(s_t) + J_perp sum_{t,i} s_t,i s_t+1,i); its mean slice energy is
    compared with exact enumeration over all 2^(N*P) configurations (N=3, P=4)."""
    import itertools
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(11)
    n, P = 3, 4
    J = rng.normal(size=(n, n))
    J = np.triu(J, 1) + np.triu(J, 1).T
    h = rng.normal(size=n) * 0.3
    beta, gamma = 1.5, 1.0
    jp = 0.5 * np.log(1.0 / np.tanh(beta * gamma / P))
    orc = Oracle(ORACLE_SO)
    single = np.array(list(itertools.product([-1, 1], repeat=n)), np.int8)
    es = orc.energy(Problem.dense(h, J), single)
    idx = np.array(list(itertools.product(range(2 ** n), repeat=P)))
    E = es[idx]
    sp = single[idx].astype(float)
    B = (sp * np.roll(sp, -1, axis=1)).sum(axis=(1, 2))
    logw = -(beta / P) * E.sum(axis=1) + jp * B
    w = np.exp(logw - logw.max())
    exact = float((w * E.mean(axis=1)).sum() / w.sum())
    dp = abi.DeviceProblem.dense(ctx, h, J)
    steps = 4000
    r = abi.sqa_run(dp, np.full(steps, beta), np.full(steps, gamma), P, 64, 1, 0, 0, abi.MODE_FP32, 5)
    est = float(np.mean(r["energy_trace"][200:]))
    print("stationary exact", exact, "fp32", est, flush=True)
    assert abs(est - exact) < 0.02 * (1 + abs(exact))
