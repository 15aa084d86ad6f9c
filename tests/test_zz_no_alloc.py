"""qa_sqa_run_noalloc: the FP32 SQA run that allocates nothing (its buffers live in a static
per-device scratch array) returns exactly what qa_sqa_run returns in FP32 mode -- best state,
FP64 best energy, energy trace and acceptance -- on two calls in a row, the second smaller than the
first (the scratch keeps the first call's contents, which must not leak into the second), for the
pipelined kernel, the two-barrier kernel with worldline and cluster sweeps, and a state too large
for LDS. A run larger than the scratch is refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from qlatanneal_amd.abi import Context
    return Context(0)


def _sk(rng, n):
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def _same(a, b):
    assert np.array_equal(a["best_state"], b["best_state"])
    assert a["best_energy"] == b["best_energy"]
    assert np.array_equal(a["energy_trace"], b["energy_trace"])


def test_noalloc_equals_sqa_run_two_calls(ctx):
    from qlatanneal_amd import abi
    rng = np.random.default_rng(77)
    # (n, P, R, steps, spb, wl, cl): pipelined slice sweeps; then a smaller two-barrier run with
    # worldline and cluster sweeps; then a state kept in global memory (n * P too large for LDS)
    cases = [(300, 16, 64, 12, 2, 0, 0), (70, 8, 12, 9, 1, 1, 1), (1500, 32, 2, 3, 1, 0, 0)]
    for (n, P, R, steps, spb, wl, cl), seed in zip(cases, (11, 12, 13)):
        print("noalloc case", n, P, R, steps, spb, wl, cl, flush=True)
        dp = abi.DeviceProblem.dense(ctx, rng.normal(0, 0.1, n), _sk(rng, n), 0.25)
        betas, gammas = np.geomspace(0.2, 3.0, steps), np.geomspace(3.0, 0.02, steps)
        ref = abi.sqa_run(dp, betas, gammas, P, R, spb, wl, cl, abi.MODE_FP32, seed)
        got = abi.sqa_run_noalloc(dp, betas, gammas, P, R, spb, wl, cl, seed)
        _same(got, ref)
        assert got["best_energy"] == dp.energy(got["best_state"])[0]
        if wl == 0 and cl == 0:
            assert got["accept_rate"] == ref["accept_rate"]


def test_noalloc_refuses_oversized_run(ctx):
    from qlatanneal_amd import abi
    n = 4096
    dp = abi.DeviceProblem.sparse(ctx, np.zeros(n), [(i, i + 1, 1.0) for i in range(n - 1)])
    with pytest.raises(abi.QaInvalidArgument, match="scratch"):
        abi.sqa_run_noalloc(dp, [1.0], [1.0], 64, 64, 1)  # 64 * 64 * 4096 spins alone exceed 64 MiB
