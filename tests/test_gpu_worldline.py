"""Pipelined worldline sweeps (sqa_pipe_wl_kernel) against the two-barrier kernel.

The worldline move (sqa_annealer.cpp:294-315) flips site i in every slice of a replica together,
on beta/P times the summed classical change. The pipelined kernel keeps the visit orders, Philox
words and FP32 arithmetic of the two-barrier kernel (slice sums in slice order, the same fmaf per
field element in the same order), so both must return the same samples bit for bit;
QANNEAL_SQA_WL_PIPE=0 selects the two-barrier kernel. It runs when a launch holds only worldline
sweeps on an LDS-resident dense state: the second launch of a split step (slice sweeps in the
pipelined slice kernel, then worldline sweeps and the step tail) and worldline-only runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from qlatanneal_amd.abi import Context
    return Context(0)


def _sk(rng, n):
    J = np.triu(rng.normal(0, 1 / np.sqrt(n), (n, n)), 1)
    return J + J.T


KEYS = ("best_state", "best_energy", "energy_trace")


# (n, P, replicas, slice sweeps, worldline sweeps, steps): the split path (C4's P=16, 64 replicas),
# worldline-only runs, a row that ends inside a 4-float group, P = 5 / 32 / 64 (12, 2 and 1
# replicas per group), a row of 150 groups (4 per lane)
CASES = [(300, 16, 64, 4, 2, 3), (130, 8, 24, 0, 3, 4), (257, 16, 8, 3, 2, 3), (200, 5, 30, 2, 2, 3),
         (96, 32, 4, 2, 2, 2), (64, 64, 2, 1, 2, 2), (600, 16, 4, 2, 2, 2), (40, 16, 3, 0, 4, 2),
         # no slice ring beside the state: slice sweeps in the two-barrier kernel, worldline sweeps in
         # the pipelined kernel on a ring of a few rows
         (1100, 16, 16, 2, 2, 2), (1465, 16, 4, 2, 2, 2)]


@pytest.mark.parametrize("n,P,R,spb,wl,steps", CASES)
def test_pipelined_worldline_sweeps_bit_identical(ctx, monkeypatch, n, P, R, spb, wl, steps):
    from qlatanneal_amd import abi
    print(f"case n={n} P={P} R={R} spb={spb} wl={wl} steps={steps}", flush=True)
    rng = np.random.default_rng(1000 + n + P)
    dp = abi.DeviceProblem.dense(ctx, rng.normal(0, 0.2, n), _sk(rng, n), 0.5)
    betas, gammas = np.linspace(0.5, 3.0, steps), np.geomspace(2.0, 0.05, steps)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("QANNEAL_SQA_WL_PIPE", mode)
        out[mode] = abi.sqa_run(dp, betas, gammas, P, R, spb, wl, 0, abi.MODE_FP32, 77)
    for k in KEYS:
        assert np.array_equal(out["0"][k], out["1"][k]), k
    assert out["1"]["best_energy"] == dp.energy(out["1"]["best_state"])[0]


def test_pipelined_worldline_sweeps_parallel_tempering(ctx, monkeypatch):
    """Worldline-only SQA-PT steps: every replica at its rung's beta/P (per-replica parameters)."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(7)
    n, P = 256, 16
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(rng, n))
    b, g = np.geomspace(0.5, 4.0, 6), np.geomspace(2.0, 0.05, 6)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("QANNEAL_SQA_WL_PIPE", mode)
        out[mode] = abi.sqapt_run(dp, b, g, P, 0, 3, 5, 1, 0, ladders=3, seed=9)
    for k in ("final_states", "final_energies", "best_state", "best_energy", "average_energy_trace",
              "swap_acceptance_trace"):
        assert np.array_equal(out["0"][k], out["1"][k]), k
