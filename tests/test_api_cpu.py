"""The drop-in API on the CPU side: the models' host FP64 methods and QUBO conversion against the
oracle, schedule helpers against their closed forms, solve()'s argument errors, and the
refusal to run a sweep without the GPU backend."""
import math

import numpy as np
import pytest

qa = pytest.importorskip("qlatanneal_amd")
from oracle.orc import Oracle, Problem, ORACLE_SO  # noqa: E402
from tests.golden_io import load, problem, dec, bits_equal  # noqa: E402


@pytest.fixture(scope="module")
def orc():
    return Oracle(ORACLE_SO)


def _api_model(d):
    p = problem(d)
    if p.kind == 0:
        return qa.DenseIsing(p.h, p.J, p.c)
    if p.kind == 1:
        return qa.SparseIsing(p.h, [qa.SparseEdge(i, j, v) for i, j, v in p.edges], p.n, p.c)
    return qa.HigherOrderIsing([(list(v), c) for v, c in p.terms], p.n)


def test_models_match_golden(orc):
    for c in load("hamiltonian.json")["cases"]:
        m = _api_model(c["problem"])
        p = problem(c["problem"])
        spins = np.array(c["spins"], np.int8)
        if p.kind == 2 and p.c != 0.0:
            continue  # HigherOrderIsing's Python ctor takes no constant (reference bindings.cpp:282-301)
        e = np.array([m.energy(s.tolist()) for s in spins])
        assert bits_equal(e, dec(c["energy"]))
        assert bits_equal([m.delta_energy(spins[0].tolist(), c["flip"])], dec(c["delta"]))


def test_qubo_to_ising_matches_oracle(orc):
    rng = np.random.default_rng(4)
    Q = rng.integers(-50, 51, (9, 9)).astype(float)
    d = qa.QUBO(Q).to_ising()
    h, J, c = orc.qubo_to_ising(Q)
    assert bits_equal(d.fields(), h) and bits_equal(d.couplings().ravel(), J.ravel())
    assert d.constant() == c
    entries = [(0, 1, 2.0), (1, 2, -1.0), (2, 2, 3.0)]
    d2 = qa.QUBO(entries, 3).to_ising()
    Qd = np.zeros((3, 3))
    for i, j, v in entries:
        Qd[i, j] += v
    h2, J2, c2 = orc.qubo_to_ising(Qd)
    assert bits_equal(d2.fields(), h2) and d2.constant() == c2


def test_hubo_consolidation_energy_matches_oracle(orc):
    rng = np.random.default_rng(9)
    n = 8
    terms = [((i,), float(rng.normal())) for i in range(n)]
    terms += [(tuple(int(x) for x in rng.choice(n, 3, replace=False)), float(rng.normal())) for _ in range(12)]
    terms += [((1, 1, 4), 0.5), ((2, 2), 0.25), ((4, 1), -0.5)]
    m = qa.HigherOrderIsing({t: c for t, c in terms} if False else dict((tuple(t), c) for t, c in terms), n)
    ref = Problem.hubo(n, list(dict((tuple(t), c) for t, c in terms).items()))
    for _ in range(10):
        s = rng.choice([-1, 1], n).astype(np.int8)
        assert m.energy(s.tolist()) == orc.energy(ref, s)[0]


def test_schedule_helpers_closed_forms():
    s = qa.auto_schedule_sqa()
    assert np.allclose(s.betas, np.linspace(0.1, 4.0, 50)) and np.allclose(s.gammas, np.geomspace(5.0, 0.01, 50))
    a = qa.AnnealSchedule.linear(0.5, 2.5, 5)
    assert list(a.betas) == [0.5, 1.0, 1.5, 2.0, 2.5]
    assert qa.AnnealSchedule.linear(0.1, 3.0, 1).betas == [3.0]
    with pytest.raises(ValueError):
        qa.AnnealSchedule.linear(0.1, 3.0, 0)
    with pytest.raises(ValueError):
        qa.SQASchedule.from_vectors([1.0, 2.0], [1.0])
    for b, g, m in [(2.0, 3.0, 32), (32.0, 1e-3, 32), (1.0, 1e-20, 8)]:
        x = max(b * g / m, 1e-12)
        assert qa.j_perp_from_beta_gamma(b, g, m) == pytest.approx(0.5 * math.log(1.0 / math.tanh(x)))


def _expected_jrms(J):
    off = J[np.triu_indices(J.shape[0], k=1)]
    nz = off[off != 0.0]
    return math.sqrt(float(np.mean(nz * nz)))


def test_j_rms_matches_analytic():
    # tests/test_optimal_schedule_resolution.py:86-90 of the reference
    rng = np.random.default_rng(2)
    n = 40
    J = np.triu(rng.normal(0, 1 / math.sqrt(n), (n, n)), 1)
    J = J + J.T
    assert qa.j_rms_from_problem(J) == pytest.approx(_expected_jrms(J), rel=1e-9)
    assert qa.j_rms_from_problem(qa.DenseIsing(np.zeros(n), J)) == pytest.approx(_expected_jrms(J), rel=1e-9)
    assert qa.j_rms_from_problem({(0, 1): 3.0, (1, 2): 4.0}) == pytest.approx(math.sqrt(12.5))


def test_solve_argument_errors():
    Q = np.eye(3)
    with pytest.raises(ValueError):
        qa.solve(Q, method="nope")
    with pytest.raises(ValueError):
        qa.solve(Q, method="ctpimc")
    with pytest.raises(ValueError):
        qa.solve(Q, schedule_type="bogus")
    with pytest.raises(ValueError):
        qa.solve(qa.DenseIsing(np.zeros(2), np.zeros((2, 2))), return_bits=True)
    with pytest.raises(ValueError):
        qa.solve(Q, method="sqa_chi", schedule_type="optimal")


def test_cpu_backend_refuses_to_run():
    d = qa.DenseIsing(np.zeros(3), np.zeros((3, 3)))
    a = qa.SQAAnnealer(d, qa.auto_schedule_sqa(5), 4, 1, backend="cpu")
    with pytest.raises(RuntimeError, match="GPU only"):
        a.run(1, 0)
    with pytest.raises(ValueError):
        qa.SQAAnnealer(d, qa.auto_schedule_sqa(5), 4, 1, backend="tpu")
    with pytest.raises(RuntimeError):
        qa.Annealer(d, qa.auto_schedule_sa(), backend="cuda")
    r = qa.ReplicaAnnealer(d, qa.AnnealSchedule.linear(0.1, 1.0, 4), 2, backend="cpu")
    with pytest.raises(RuntimeError, match="GPU only"):
        r.run(1)
    t = qa.ParallelTemperingAnnealer(d, [0.5, 1.0], backend="cpu")
    with pytest.raises(RuntimeError, match="GPU only"):
        t.run(1, 2)


def test_classical_argument_errors():
    """The reference's std::invalid_argument cases (replica_annealer.cpp:17-21,47-49,
    parallel_tempering.cpp:16-18,39-47) raise ValueError before any device work."""
    d = qa.DenseIsing(np.zeros(3), np.zeros((3, 3)))
    with pytest.raises(ValueError, match="replicas"):
        qa.ReplicaAnnealer(d, qa.AnnealSchedule.linear(0.1, 1.0, 4), 0, backend="cpu")
    with pytest.raises(ValueError, match="at least one beta"):
        qa.ReplicaAnnealer(d, qa.AnnealSchedule(), 2, backend="cpu")
    with pytest.raises(ValueError, match="sweeps_per_beta"):
        qa.ReplicaAnnealer(d, qa.AnnealSchedule.linear(0.1, 1.0, 4), 2, backend="cpu").run(0)
    with pytest.raises(ValueError, match="two betas"):
        qa.ParallelTemperingAnnealer(d, [1.0], backend="cpu")
    t = qa.ParallelTemperingAnnealer(d, [0.5, 1.0], backend="cpu")
    for args, msg in (((0, 1, 1), "sweeps_per_step"), ((1, 0, 1), "steps"), ((1, 1, 0), "swap_interval")):
        with pytest.raises(ValueError, match=msg):
            t.run(*args)


def test_chi_annealed_ladder_layout():
    """Per-step SQA-PT ladders from a run_chi schedule: the coldest rung anneals on the schedule,
    the hottest trails it by beta_ratio, rungs between are geometric, Gamma is shared per step."""
    from qlatanneal_amd.surrogate import chi_annealed_ladder
    b = np.array([0.1, 1.0, 4.0, 16.0, 16.0])
    g = np.array([3.0, 3.0, 1.0, 0.1, 0.001])
    sb, sg = chi_annealed_ladder(b, g, 3, 0.25)
    assert sb.shape == sg.shape == (5, 3)
    assert np.array_equal(sb[:, 2], b) and np.allclose(sb[:, 0], 0.25 * b) and np.allclose(sb[:, 1], 0.5 * b)
    assert np.all(sg == g[:, None])
    sb1, _ = chi_annealed_ladder(b, g, 4, 1.0)
    assert np.all(sb1 == b[:, None])
    for bad in (dict(rungs=1), dict(beta_ratio=0.0), dict(beta_ratio=1.5)):
        kw = dict(rungs=3, beta_ratio=0.5)
        kw.update(bad)
        with pytest.raises(ValueError):
            chi_annealed_ladder(b, g, **kw)
    with pytest.raises(ValueError):
        chi_annealed_ladder(b, g[:-1], 3)
    # gamma_max: the ladder starts at the first step with Gamma <= gamma_max
    sb2, sg2 = chi_annealed_ladder(b, g, 3, 0.25, gamma_max=1.0)
    assert sb2.shape == (3, 3) and np.array_equal(sb2, sb[2:]) and np.array_equal(sg2, sg[2:])
    with pytest.raises(ValueError):
        chi_annealed_ladder(b, g, 3, gamma_max=1e-4)
