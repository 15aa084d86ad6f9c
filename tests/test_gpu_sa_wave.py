"""The small-problem SA kernel (one wavefront per read, dense n <= 128) against the block kernel
(QANNEAL_SA_KERNEL=block): both consume the same Philox words with the same FP32 arithmetic, so
the spin trajectories (final states, magnetization traces) must be identical; running energies may
differ only by the rounding of the FP64 start energy (a different summation order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sk(rng, n):
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def test_sa_wave_kernel_equals_block_kernel(monkeypatch):
    from qlatanneal_amd import abi
    ctx = abi.Context(0)
    rng = np.random.default_rng(61)
    for n in (3, 17, 64, 65, 100, 128):
        J, h = _sk(rng, n), rng.normal(0, 0.2, n)
        dp = abi.DeviceProblem.dense(ctx, h, J, 0.25)
        betas = np.geomspace(0.1, 3.0, 20)
        for reads in (1, 5, 33):
            print("n", n, "reads", reads, flush=True)
            out = []
            for kern in (None, "block"):
                if kern:
                    monkeypatch.setenv("QANNEAL_SA_KERNEL", kern)
                else:
                    monkeypatch.delenv("QANNEAL_SA_KERNEL", raising=False)
                out.append(abi.sa_run(dp, betas, 3, reads, abi.MODE_FP32, seed=7 + n))
            a, b = out
            assert np.array_equal(a["final_states"], b["final_states"]), (n, reads)
            assert np.array_equal(a["magnetization_traces"], b["magnetization_traces"])
            np.testing.assert_allclose(a["energy_traces"], b["energy_traces"], rtol=0, atol=1e-9)
            np.testing.assert_allclose(a["best_energies"], b["best_energies"], rtol=0, atol=1e-9)
            # chained-run inputs: start states, start energies, per-read betas, stream offset
            init = a["final_states"]
            e0 = a["final_energies"]
            rb = np.linspace(0.5, 2.0, reads)
            out = []
            for kern in (None, "block"):
                if kern:
                    monkeypatch.setenv("QANNEAL_SA_KERNEL", kern)
                else:
                    monkeypatch.delenv("QANNEAL_SA_KERNEL", raising=False)
                out.append(abi.sa_run(dp, betas[:4], 2, reads, abi.MODE_FP32, seed=3, read_betas=rb,
                                      init_states=init, init_energies=e0, counter_base=9))
            a2, b2 = out
            assert np.array_equal(a2["final_states"], b2["final_states"])
            assert np.array_equal(a2["energy_traces"], b2["energy_traces"])  # same start energy: exact
    monkeypatch.delenv("QANNEAL_SA_KERNEL", raising=False)


def test_sa_fp32_distribution_vs_oracle_256_reads():
    """Residual-energy distributions of FP32 SA against the oracle (the reference's CPU
    algorithm), two-sample KS p > 0.01 over 256 reads: the C1 configuration (N=64 SK, geometric
    beta 0.1 -> 3.0, 100 x 10, wave kernel), a short anneal whose energies are broad (wave kernel)
    and a larger instance (N=200, block kernel). Energies are rounded to 1e-9 so exact ties stay
    ties."""
    from scipy.stats import ks_2samp
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    ctx = abi.Context(0)
    orc = Oracle(ORACLE_SO)
    rng = np.random.default_rng(71)
    geo = 0.1 * 30.0 ** (np.arange(100) / 99.0)
    for n, betas, spb in [(64, geo, 10), (64, np.linspace(0.1, 1.5, 12), 1), (200, np.linspace(0.1, 2.0, 20), 2)]:
        J = _sk(rng, n)
        dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
        gpu = abi.sa_run(dp, betas, spb, 256, abi.MODE_FP32, seed=5 + n)["best_energies"]
        p = Problem.dense(np.zeros(n), J)
        cpu = np.array([orc.sa_run(p, betas, spb, 1000 + k)["best_energy"] for k in range(256)])
        pv = ks_2samp(np.round(gpu, 9), np.round(cpu, 9)).pvalue
        print("SA KS", n, len(betas), spb, "p", pv, "gpu mean", gpu.mean(), "cpu mean", cpu.mean(), flush=True)
        assert pv > 0.01
