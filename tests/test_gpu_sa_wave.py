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
