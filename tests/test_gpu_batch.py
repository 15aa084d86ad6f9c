"""Batched SQA (qa_sqa_batch_run): many uneven problems in one persistent launch.

The oracle for batching is the unbatched run: with autotune off, every instance of the batch must
return exactly what its own qa_sqa_run call returns (same replica groups, same Philox counters,
same sweep kernels), whichever sweep variant the instance lands on (pipelined dense, two-barrier
with the state in LDS, HUBO, worldline sweeps). The hipGraph-captured launch must equal the direct
launch. With the autotuner on, replica groups change, so results are checked for exact FP64
sample energies only.
"""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="the persistent batch launch is disabled: its first GPU run did not "
                                      "complete and the cause has not been shown (qa_sqa_batch_run refuses "
                                      "to launch until it is)")]


@pytest.fixture(scope="module")
def ctx():
    from qlatanneal_amd.abi import Context
    return Context(0)


def _qubo_ising(rng, n, density):
    """Integer QUBO Q_ij ~ U{-100..100} on a random pair mask, as an Ising (h, J, c) by
    QUBO::to_ising's rule (qubo.cpp:33-69): W = (Q + Q^T)/2, h = 1/2 sum W, J = W/2 off the
    diagonal, c = 1/4 (tr W + sum W)."""
    Q = rng.integers(-100, 101, (n, n)).astype(np.float64)
    Q = np.triu(Q)
    mask = np.triu(rng.random((n, n)) < density, 1) | np.eye(n, dtype=bool)
    Q = np.where(mask, Q, 0.0)
    W = (Q + Q.T) / 2.0
    h = 0.5 * W.sum(axis=1)
    J = W / 2.0
    np.fill_diagonal(J, 0.0)
    c = 0.25 * (np.trace(W) + W.sum())
    return h, J, c


def _hubo(rng, n):
    """3-body HUBO: 4n distinct triplets with N(0,1) coefficients, plus N(0,1) fields."""
    seen, terms = set(), []
    while len(terms) < 4 * n:
        t = tuple(sorted(rng.choice(n, 3, replace=False).tolist()))
        if t not in seen:
            seen.add(t)
            terms.append(t)
    off = np.arange(0, 3 * len(terms) + 1, 3)
    var = np.array([v for t in terms for v in t], np.uint32)
    co = rng.normal(0, 1, len(terms))
    return rng.normal(0, 1, n), off, var, co


def _instances(ctx, rng):
    from qlatanneal_amd import abi
    probs, kinds = [], []
    for n, dens in [(50, 1.0), (130, 0.3), (257, 0.05), (600, 0.7), (64, 0.5), (301, 1.0)]:
        h, J, c = _qubo_ising(rng, n, dens)
        probs.append(abi.DeviceProblem.dense(ctx, h, J, c))
        kinds.append(("qubo", n))
    for n in (40, 96):
        h, off, var, co = _hubo(rng, n)
        probs.append(abi.DeviceProblem.hubo(ctx, n, h, off, var, co, 0.0))
        kinds.append(("hubo", n))
    return probs, kinds


@pytest.mark.parametrize("wl", [0, 1])
def test_batch_equals_unbatched_runs(ctx, wl):
    from qlatanneal_amd import abi
    rng = np.random.default_rng(2024 + wl)
    probs, kinds = _instances(ctx, rng)
    P, steps = 16, 6
    betas, gammas = np.geomspace(0.05, 2.0, steps), np.geomspace(3.0, 0.05, steps)
    reads = rng.integers(2, 9, len(probs))
    seeds = rng.integers(1, 2**62, len(probs))
    b = abi.sqa_batch_run(probs, betas, gammas, P, 2, reads, seeds, worldline_sweeps=wl)
    g = abi.sqa_batch_run(probs, betas, gammas, P, 2, reads, seeds, worldline_sweeps=wl, use_graph=True)
    assert b["attempts"] == sum(steps * (2 + wl) * int(r) * P * p.n for r, p in zip(reads, probs))
    for k, p in enumerate(probs):
        print("batch instance", k, kinds[k], "reads", reads[k], flush=True)
        one = abi.sqa_run(p, betas, gammas, P, int(reads[k]), 2, wl, 0, abi.MODE_FP32, int(seeds[k]))
        assert np.array_equal(b["best_states"][k], one["best_state"]), k
        assert b["best_energies"][k] == one["best_energy"], k
        assert np.array_equal(g["best_states"][k], one["best_state"]), k
        assert g["best_energies"][k] == one["best_energy"], k


def test_batch_autotuned_samples_exact(ctx):
    from qlatanneal_amd import abi
    rng = np.random.default_rng(77)
    probs, kinds = _instances(ctx, rng)
    P, steps = 16, 5
    betas, gammas = np.geomspace(0.05, 2.0, steps), np.geomspace(3.0, 0.05, steps)
    reads = np.full(len(probs), 16)
    seeds = rng.integers(1, 2**62, len(probs))
    r = abi.sqa_batch_run(probs, betas, gammas, P, 2, reads, seeds, autotune=True)
    for k, p in enumerate(probs):
        assert r["best_energies"][k] == p.energy(r["best_states"][k])[0], (k, kinds[k])
