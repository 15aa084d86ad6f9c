"""Engines driven from several host threads. Each HIP backend is bound to one of the library's
contexts (one HIP stream each), so runs of different engines overlap on the GPU; every run must
equal the same run made alone, bit for bit, in both precisions."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

qa = pytest.importorskip("qlatanneal_amd")


def _models(rng):
    out = []
    for k in range(5):
        n = 40 + 8 * k
        J = np.triu(rng.normal(0, 1 / np.sqrt(n), (n, n)), 1)
        out.append(qa.DenseIsing(rng.normal(0, 0.2, n), J + J.T))
    n = 60
    out.append(qa.SparseIsing(np.zeros(n), [qa.SparseEdge(i, (i + 1) % n, float(rng.normal())) for i in range(n)]
                              + [qa.SparseEdge(i, (i + 7) % n, float(rng.normal())) for i in range(n)], n))
    out.append(qa.HigherOrderIsing([([i, (i + 1) % 30, (i + 3) % 30], float(rng.normal())) for i in range(30)]
                                   + [([i, (i + 2) % 30], float(rng.normal())) for i in range(30)], 30))
    return out


def _sqa(m, backend, seed):
    betas = np.linspace(0.3, 3.0, 10).tolist()
    gammas = np.geomspace(3.0, 0.02, 10).tolist()
    a = qa.SQAAnnealer(m, qa.SQASchedule.from_vectors(betas, gammas), 8, 4, backend=backend)
    return a, seed


def _run_sqa(pair):
    a, seed = pair
    a.set_seed(seed)
    r = a.run(3, 1)
    return r.best_energy, list(r.best_state.spins), list(r.energy_trace)


def _run_sa(pair):
    a, seed = pair
    res = a.run_reads(4, 6, seed)
    return [(r.best_energy, list(r.best_state.spins)) for r in res]


@pytest.mark.parametrize("backend", ("hip", "hip_fp64"))
def test_threaded_engines_equal_sequential(backend):
    rng = np.random.default_rng(61)
    models = _models(rng)
    print("threaded engines", backend, len(models), flush=True)
    alone = [_run_sqa(_sqa(m, backend, 200 + i)) for i, m in enumerate(models)]
    pairs = [_sqa(m, backend, 200 + i) for i, m in enumerate(models)]
    with ThreadPoolExecutor(8) as ex:  # more threads than contexts: runs wait for a free one
        together = list(ex.map(_run_sqa, pairs))
    assert together == alone
    # a second round on the same engines: runs may now land on other contexts than the first time
    # (the problem is copied into whichever context is free), with the same results
    with ThreadPoolExecutor(3) as ex:
        again = list(ex.map(_run_sqa, pairs[::-1]))
    assert again[::-1] == alone
    sched = qa.AnnealSchedule.linear(0.2, 3.0, 16)
    alone = [_run_sa((qa.Annealer(m, sched, backend=backend), 30 + i)) for i, m in enumerate(models)]
    with ThreadPoolExecutor(4) as ex:
        together = list(ex.map(_run_sa, [(qa.Annealer(m, sched, backend=backend), 30 + i)
                                         for i, m in enumerate(models)]))
    assert together == alone


def test_solve_concurrent_reads_equal_sequential_loop():
    """solve() with a seed spreads reads over engines on several HIP contexts; read r must equal
    the single read of solve(..., reads=1, seed=seed + r), which runs set_seed(seed + r) and one
    run on one engine (the reference's read loop, solver.py:679)."""
    assert qa.hip_context_count() >= 1
    rng = np.random.default_rng(62)
    n = 24
    J = np.triu(rng.normal(0, 1 / np.sqrt(n), (n, n)), 1)
    d = qa.DenseIsing(rng.normal(0, 0.2, n), J + J.T)
    sched = qa.SQASchedule.from_vectors(np.linspace(0.5, 3.0, 8).tolist(), np.geomspace(3.0, 0.05, 8).tolist())
    cases = [("sqa", 7, dict(schedule=sched, trotter_slices=8, replicas=2, sweeps_per_beta=3)),
             ("sqapt", 5, dict(schedule=sched, trotter_slices=4, pt_steps=4, sweeps_per_beta=2)),
             ("sqa_chi", 3, dict(trotter_slices=8, replicas=2, chi_scan_points=4, chi_scan_sweeps=3,
                                 chi_scan_burn=1, sweeps_per_beta=2, optimal_num_steps=12))]
    for backend in ("hip", "hip_fp64"):
        for method, reads, kw in cases:
            print("solve concurrent reads", backend, method, flush=True)
            many = qa.solve(d, method=method, reads=reads, seed=40, backend=backend, progress=False, **kw)
            for r in range(reads):
                one = qa.solve(d, method=method, reads=1, seed=40 + r, backend=backend, progress=False, **kw)
                assert many.energies[r] == one.energies[0], (method, r)
                assert list(many.samples[r]) == list(one.samples[0]), (method, r)
