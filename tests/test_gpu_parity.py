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


# sweep implementations selected by an environment switch: the two-barrier lockstep sweep and the
# pipelined LDS-field sweep (the register-field sweep is switched off in the library)
_SWEEP_PATHS = {"lockstep": {"QANNEAL_SQA_PIPELINE": "0"},
                "pipelined": {"QANNEAL_SQA_PIPELINE": "1"}}


def test_sqa_fp32_global_state_update_paths_bit_identical(ctx, monkeypatch):
    """Global-memory state (fields too large for LDS): the float4 batched field update (all loads of
    a batch of flipped chains issued before its stores) must reproduce the column-loop update bit
    for bit -- slice, cluster and worldline sweeps, SQA-PT, and the sqa_chi checkerboard."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(43)
    for n, P, R, spb, wl, cl in [(1100, 64, 2, 1, 1, 1), (2048, 64, 1, 2, 0, 0), (700, 48, 3, 1, 0, 1)]:
        print("global-state case", n, P, R, spb, wl, cl, flush=True)
        J = _sk(rng, n)
        dp = abi.DeviceProblem.dense(ctx, rng.normal(0, 0.2, n), J, 0.5)
        betas, gammas = np.full(4, 2.0), np.geomspace(3.0, 0.05, 4)
        res = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_SQA_GLOBAL_F4", v)
            res.append(abi.sqa_run(dp, betas, gammas, P, R, spb, wl, cl, abi.MODE_FP32, 17))
        a, b = res
        assert a["accept_rate"] == b["accept_rate"]
        assert np.array_equal(a["best_state"], b["best_state"]) and bits_equal([a["best_energy"]], [b["best_energy"]])
        assert bits_equal(a["energy_trace"], b["energy_trace"])
    n, P = 1024, 64
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(rng, n))
    res = []
    for v in ("1", "0"):
        monkeypatch.setenv("QANNEAL_SQA_GLOBAL_F4", v)
        res.append(abi.sqapt_run(dp, np.geomspace(0.3, 3.0, 3), np.geomspace(3.0, 0.05, 3), P, 2, 0, 3, 1, 0,
                                 ladders=2, mode=abi.MODE_FP32, seed=5))
    for key in ("final_states", "best_state"):
        assert np.array_equal(res[0][key], res[1][key]), key
    for key in ("final_energies", "best_energy", "average_energy_trace", "swap_acceptance_trace"):
        assert bits_equal(res[0][key].ravel(), res[1][key].ravel()), key
    res = []
    for v in ("1", "0"):
        monkeypatch.setenv("QANNEAL_SQA_GLOBAL_F4", v)
        res.append(abi.sqa_chi_run(dp, 64, 1, 2.0, 3.0, 0.05, num_steps=6, sweeps_per_step=2, scan_points=4,
                                   scan_sweeps=2, scan_burn=1, seed=9))
    assert np.array_equal(res[0]["best_state"], res[1]["best_state"])
    assert bits_equal(res[0]["energy_trace"], res[1]["energy_trace"])
    monkeypatch.delenv("QANNEAL_SQA_GLOBAL_F4", raising=False)


def test_sqa_fp32_global_state_prefetch_bit_identical(ctx, monkeypatch):
    """The global-memory slice sweep loads each chain's next decision inputs a step ahead (own flip
    added from J[i_k][i_k+1], the previous slice's spin taken from its lane's decision). With the
    state forced into global memory it must reproduce the LDS-state kernels and the column-loop
    global path bit for bit, for P = 1, 2, 3 (ring closure onto the lane itself, onto the lane
    before, and the general case), several replicas per workgroup, and SQA-PT."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(44)
    for n, P, R in [(64, 1, 5), (96, 2, 3), (132, 3, 4), (200, 8, 9), (256, 16, 4)]:
        print("prefetch case", n, P, R, flush=True)
        dp = abi.DeviceProblem.dense(ctx, rng.normal(0, 0.2, n), _sk(rng, n), 0.25)
        betas, gammas = np.geomspace(0.3, 3.0, 5), np.geomspace(3.0, 0.05, 5)
        res = []
        for state, f4, sl in (("lds", "1", "1"), ("global", "1", "1"), ("global", "1", "0"), ("global", "0", "1")):
            monkeypatch.setenv("QANNEAL_SQA_STATE", state)
            monkeypatch.setenv("QANNEAL_SQA_GLOBAL_F4", f4)
            monkeypatch.setenv("QANNEAL_SQA_SPINS_LDS", sl)
            res.append(abi.sqa_run(dp, betas, gammas, P, R, 2, 0, 0, abi.MODE_FP32, 23))
        for r in res[1:]:
            assert r["accept_rate"] == res[0]["accept_rate"], (n, P, R)
            assert np.array_equal(r["best_state"], res[0]["best_state"]), (n, P, R)
            assert bits_equal([r["best_energy"]], [res[0]["best_energy"]]), (n, P, R)
            assert bits_equal(r["energy_trace"], res[0]["energy_trace"]), (n, P, R)
    n, P = 512, 32
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(rng, n))
    res = []
    for state, sl in (("lds", "1"), ("global", "1"), ("global", "0")):
        monkeypatch.setenv("QANNEAL_SQA_STATE", state)
        monkeypatch.setenv("QANNEAL_SQA_GLOBAL_F4", "1")
        monkeypatch.setenv("QANNEAL_SQA_SPINS_LDS", sl)
        res.append(abi.sqapt_run(dp, np.geomspace(0.3, 3.0, 4), np.geomspace(3.0, 0.05, 4), P, 3, 0, 4, 1, 0,
                                 ladders=3, mode=abi.MODE_FP32, seed=6))
    for r in res[1:]:
        for key in ("final_states", "best_state"):
            assert np.array_equal(res[0][key], r[key]), key
        for key in ("final_energies", "best_energy", "average_energy_trace", "swap_acceptance_trace"):
            assert bits_equal(res[0][key].ravel(), r[key].ravel()), key
        assert 0.0 < r["accept_rate"] < 1.0 and r["accept_rate"] == res[0]["accept_rate"]
    monkeypatch.delenv("QANNEAL_SQA_STATE", raising=False)
    monkeypatch.delenv("QANNEAL_SQA_GLOBAL_F4", raising=False)
    monkeypatch.delenv("QANNEAL_SQA_SPINS_LDS", raising=False)


def test_sqa_fp32_lookahead_sweep_bit_identical(ctx, monkeypatch):
    """Global fields with the spins in LDS: the look-ahead slice sweep (decisions from LDS, field
    elements and couplings loaded three steps ahead by the helper wavefront, the chain's own last
    three flips added by the decision) against the one-step-prefetch sweep and the LDS-state kernel,
    bit for bit: row lengths that end inside a wavefront's share (n = 300, 1100), one to 64 chains
    per group, worldline and cluster moves between the slice sweeps, and SQA-PT at P = 64."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(45)
    env = ("QANNEAL_SQA_STATE", "QANNEAL_SQA_GLOBAL_F4", "QANNEAL_SQA_SPINS_LDS", "QANNEAL_SQA_LOOKAHEAD")
    configs = (("lds", "1", "1", "1"), ("global", "1", "1", "0"), ("global", "1", "1", "1"))
    for n, P, R, wl, cl in [(64, 1, 5, 0, 0), (132, 3, 4, 0, 0), (300, 16, 8, 0, 0), (256, 64, 2, 1, 1),
                            (1100, 8, 3, 0, 0), (64, 16, 1024, 0, 0), (2048, 64, 2, 0, 0)]:
        print("lookahead case", n, P, R, wl, cl, flush=True)
        dp = abi.DeviceProblem.dense(ctx, rng.normal(0, 0.2, n), _sk(rng, n), 0.25)
        betas, gammas = np.geomspace(0.3, 3.0, 4), np.geomspace(3.0, 0.05, 4)
        res = []
        for cfg in configs:
            for k, v in zip(env, cfg):
                monkeypatch.setenv(k, v)
            res.append(abi.sqa_run(dp, betas, gammas, P, R, 2, wl, cl, abi.MODE_FP32, 29))
        for r in res[1:]:
            assert r["accept_rate"] == res[0]["accept_rate"], (n, P, R)
            assert np.array_equal(r["best_state"], res[0]["best_state"]), (n, P, R)
            assert bits_equal([r["best_energy"]], [res[0]["best_energy"]]), (n, P, R)
            assert bits_equal(r["energy_trace"], res[0]["energy_trace"]), (n, P, R)
    n, P = 2048, 64
    print("lookahead case sqapt", n, P, flush=True)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(rng, n))
    res = []
    for cfg in configs[1:]:
        for k, v in zip(env, cfg):
            monkeypatch.setenv(k, v)
        res.append(abi.sqapt_run(dp, np.geomspace(0.3, 5.0, 8), np.geomspace(3.0, 0.05, 8), P, 10, 0, 3, 1, 0,
                                 ladders=2, mode=abi.MODE_FP32, seed=8, refresh_interval=10))
    for key in ("final_states", "best_state"):
        assert np.array_equal(res[0][key], res[1][key]), key
    for key in ("final_energies", "best_energy", "average_energy_trace", "swap_acceptance_trace", "field_drift"):
        assert bits_equal(res[0][key].ravel(), res[1][key].ravel()), key
    assert 0.0 < res[1]["accept_rate"] < 1.0 and res[1]["accept_rate"] == res[0]["accept_rate"]
    for k in env:
        monkeypatch.delenv(k, raising=False)


def test_sqa_fp32_sweep_paths_bit_identical(ctx, monkeypatch):
    """The one-barrier pipelined sweep (decisions of step t overlapped with the field updates of
    step t-1) must reproduce the two-barrier lockstep sweep bit for bit: same spins, same FP32
    fields, hence the same accept count, energy trace and best sample."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(41)
    cases = [(200, 8, 16, 1, 0, 0), (97, 5, 7, 3, 1, 1), (512, 32, 4, 2, 0, 0), (33, 1, 3, 2, 0, 0),
             (64, 64, 2, 1, 0, 1), (130, 16, 9, 2, 1, 0), (128, 12, 5, 2, 0, 0), (64, 4, 1024, 1, 0, 0),
             (301, 20, 3, 1, 0, 0), (600, 8, 2, 1, 0, 0), (1100, 4, 1, 1, 0, 0)]
    for n, P, R, spb, wl, cl in cases:
        print("sweep-path case", n, P, R, spb, wl, cl, flush=True)
        J = _sk(rng, n)
        h = rng.normal(0, 0.2, n)
        dp = abi.DeviceProblem.dense(ctx, h, J, -0.25)
        betas, gammas = np.full(12, 3.0), np.geomspace(3.0, 0.02, 12)
        res = {}
        for name, env in _SWEEP_PATHS.items():
            for k, v in env.items():
                monkeypatch.setenv(k, v)
            res[name] = abi.sqa_run(dp, betas, gammas, P, R, spb, wl, cl, abi.MODE_FP32, 99)
        a = res["lockstep"]
        for name in ("pipelined",):
            b = res[name]
            assert a["accept_rate"] == b["accept_rate"], name
            assert np.array_equal(a["best_state"], b["best_state"]), name
            assert bits_equal([a["best_energy"]], [b["best_energy"]]), name
            assert bits_equal(a["energy_trace"], b["energy_trace"]), name
    # SQA-PT runs every sweep launch through the same kernels (resume + per-chain tails)
    n, P = 48, 8
    J = _sk(rng, n)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    res = {}
    for name, env in _SWEEP_PATHS.items():
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        res[name] = abi.sqapt_run(dp, np.geomspace(0.3, 3.0, 4), np.geomspace(3.0, 0.05, 4), P, 2, 0, 6, 1, 0,
                                  ladders=8, mode=abi.MODE_FP32, seed=5)
    for name in ("pipelined",):
        for key in ("final_states", "best_state"):
            assert np.array_equal(res["lockstep"][key], res[name][key]), (name, key)
        for key in ("final_energies", "best_energy", "average_energy_trace", "swap_acceptance_trace"):
            assert bits_equal(res["lockstep"][key].ravel(), res[name][key].ravel()), (name, key)


def test_sqa_fp32_split_step_launches_bit_identical(ctx, monkeypatch):
    """Steps with cluster or worldline sweeps run their slice sweeps in the pipelined kernel and the
    rest in the two-barrier kernel (two launches per step, state passed through global memory); the
    result must equal the single launch bit for bit, with and without field refreshes."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(47)
    # ... and global-state fields (spins in LDS): slice sweeps in the look-ahead kernel, the rest in
    # the one-step-prefetch kernel (solve()'s defaults at N=2048, P=32, one replica per read)
    cases = [(200, 16, 64, 2, 1, 0, 0), (97, 5, 7, 3, 1, 1, 0), (550, 16, 64, 4, 2, 0, 0), (64, 64, 2, 1, 0, 1, 0),
             (301, 8, 9, 2, 1, 1, 6), (1465, 16, 8, 1, 1, 0, 0), (2048, 32, 1, 2, 1, 0, 0), (1024, 16, 4, 2, 1, 1, 0),
             (2048, 32, 2, 2, 1, 0, 9)]
    for n, P, R, spb, wl, cl, ri in cases:
        print("split-step case", n, P, R, spb, wl, cl, ri, flush=True)
        J = _sk(rng, n)
        dp = abi.DeviceProblem.dense(ctx, rng.normal(0, 0.2, n), J, 0.5)
        betas, gammas = np.geomspace(0.2, 3.0, 5), np.geomspace(3.0, 0.02, 5)
        res = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_SQA_SPLIT", v)
            res.append(abi.sqa_run(dp, betas, gammas, P, R, spb, wl, cl, abi.MODE_FP32, 23, refresh_interval=ri))
        a, b = res
        assert a["accept_rate"] == b["accept_rate"]
        assert np.array_equal(a["best_state"], b["best_state"]) and bits_equal([a["best_energy"]], [b["best_energy"]])
        assert bits_equal(a["energy_trace"], b["energy_trace"])
    monkeypatch.delenv("QANNEAL_SQA_SPLIT", raising=False)


def test_sqa_fp32_lists_lds_staging_bit_identical(ctx, monkeypatch):
    """HUBO incidence and term arrays and sparse adjacency staged in LDS (the two-barrier kernel's
    field updates read them there) must give the same runs, bit for bit, as reading them from
    global memory: slice, worldline and cluster sweeps, and SQA-PT."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(59)
    for n, P, R, spb, wl, cl in [(256, 16, 8, 2, 1, 0), (96, 7, 3, 1, 1, 1), (512, 32, 2, 1, 0, 0)]:
        print("hubo staging case", n, P, R, spb, wl, cl, flush=True)
        seen, terms = set(), []
        while len(terms) < 4 * n:
            t = tuple(sorted(rng.choice(n, 3, replace=False).tolist()))
            if t not in seen:
                seen.add(t)
                terms.append(t)
        off = np.arange(0, 3 * len(terms) + 1, 3)
        var = np.array([v for t in terms for v in t], np.uint32)
        dp = abi.DeviceProblem.hubo(ctx, n, rng.normal(0, 0.5, n), off, var, rng.normal(0, 1, len(terms)), 0.0)
        betas, gammas = np.geomspace(0.3, 3.0, 4), np.geomspace(3.0, 0.05, 4)
        res = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_SQA_LISTS_LDS", v)
            res.append(abi.sqa_run(dp, betas, gammas, P, R, spb, wl, cl, abi.MODE_FP32, 61))
        a, b = res
        assert a["accept_rate"] == b["accept_rate"]
        assert np.array_equal(a["best_state"], b["best_state"]) and bits_equal([a["best_energy"]], [b["best_energy"]])
        assert bits_equal(a["energy_trace"], b["energy_trace"])
        pt = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_SQA_LISTS_LDS", v)
            pt.append(abi.sqapt_run(dp, betas, gammas, P, 1, 0, 3, 1, 0, ladders=2, mode=abi.MODE_FP32, seed=7))
        for key in ("final_states", "best_state"):
            assert np.array_equal(pt[0][key], pt[1][key]), key
        for key in ("final_energies", "average_energy_trace", "swap_acceptance_trace"):
            assert bits_equal(pt[0][key].ravel(), pt[1][key].ravel()), key
    # sparse adjacency staged in LDS (a graph with repeated edges: updates stay in row order)
    for n, P, R, spb, wl, cl in [(300, 16, 4, 2, 1, 1), (2048, 32, 2, 1, 0, 0)]:
        print("sparse staging case", n, P, R, spb, wl, cl, flush=True)
        el = [(i, (i + 1) % n, float(rng.normal())) for i in range(n)]
        el += [(i, (i + 37) % n, float(rng.normal())) for i in range(0, n, 3)]
        el += el[::7]
        dp = abi.DeviceProblem.sparse(ctx, rng.normal(0, 0.3, n), el, 0.0)
        res = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_SQA_LISTS_LDS", v)
            res.append(abi.sqa_run(dp, np.geomspace(0.3, 3.0, 4), np.geomspace(3.0, 0.05, 4), P, R, spb, wl, cl,
                                   abi.MODE_FP32, 67))
        a, b = res
        assert a["accept_rate"] == b["accept_rate"]
        assert np.array_equal(a["best_state"], b["best_state"]) and bits_equal(a["energy_trace"], b["energy_trace"])
    monkeypatch.delenv("QANNEAL_SQA_LISTS_LDS", raising=False)


def test_sqa_chi_production_wave_kernel_bit_identical(ctx, monkeypatch):
    """sqa_chi pilot and production sweeps of dense and sparse problems run one wavefront per chain
    and one launch per checkerboard phase; the result must equal the two-barrier kernel's single
    launch bit for bit (pilot chi_B scan included): odd and even P (three and two colours), P = 1 and 2, n not a multiple of 4, LDS-resident
    and global-memory state in the two-barrier kernel, the largest dense size of the wave path, and
    sparse graphs with and without repeated edges."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(53)
    cases = [(12, 8, 4), (33, 5, 3), (100, 1, 2), (130, 2, 3), (257, 7, 2), (512, 32, 4), (1100, 16, 2),
             (2048, 32, 4)]
    for n, P, R in cases:
        print("chi production case", n, P, R, flush=True)
        J = _sk(rng, n)
        dp = abi.DeviceProblem.dense(ctx, rng.normal(0, 0.2, n), J, 0.25)
        kw = dict(beta=2.5, gamma_start=3.0, gamma_end=0.05, num_steps=6, sweeps_per_step=2, scan_points=4,
                  scan_sweeps=2, scan_burn=1, seed=29)
        res = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_CHI_WAVE", v)
            res.append(abi.sqa_chi_run(dp, P, R, **kw))
        a, b = res
        assert np.array_equal(a["best_state"], b["best_state"]), (n, P, R)
        assert bits_equal([a["best_energy"]], [b["best_energy"]]), (n, P, R)
        assert bits_equal(a["energy_trace"], b["energy_trace"]), (n, P, R)
        assert bits_equal(a["gamma_schedule"], b["gamma_schedule"]), (n, P, R)
        assert bits_equal(a["scan_chi_B"], b["scan_chi_B"]), (n, P, R)
        if n in (130, 1100):  # dense J rows streamed ahead of every site instead of loaded on demand
            monkeypatch.setenv("QANNEAL_CHI_WAVE", "1")
            monkeypatch.setenv("QANNEAL_CHI_ROWS", "stream")
            c = abi.sqa_chi_run(dp, P, R, **kw)
            monkeypatch.delenv("QANNEAL_CHI_ROWS")
            assert np.array_equal(a["best_state"], c["best_state"]), (n, P, R)
            assert bits_equal(a["energy_trace"], c["energy_trace"]) and bits_equal(a["scan_chi_B"], c["scan_chi_B"])
        if n in (257, 1100):  # the pilot scan alone, more points and measured sweeps
            pil = []
            for v in ("1", "0"):
                monkeypatch.setenv("QANNEAL_CHI_WAVE", v)
                pil.append(abi.chi_pilot(dp, P, R, 2.0, 3.0, 1e-3, 16, 5, 3, seed=41))
            assert bits_equal(pil[0]["scan_chi_B"], pil[1]["scan_chi_B"]), (n, P, R)
    # sparse: a 3-regular graph with unit weights (MAX-CUT, rows without repeated spins: parallel
    # neighbour updates) and a graph with repeated edges (sequential updates in row order)
    for n, P, R, dup in [(400, 32, 4, False), (257, 5, 3, False), (150, 8, 2, True), (4096, 32, 4, False)]:
        print("chi production sparse case", n, P, R, dup, flush=True)
        edges = set()
        while len(edges) < 3 * n // 2:
            perm = rng.permutation(n)
            for i in range(0, n - 1, 2):
                u, v = sorted((int(perm[i]), int(perm[i + 1])))
                if len(edges) < 3 * n // 2:
                    edges.add((u, v))
        el = [(u, v, 1.0) for u, v in sorted(edges)]
        if dup:
            el += [(u, v, float(rng.normal())) for u, v, _ in el[::5]]
        dp = abi.DeviceProblem.sparse(ctx, rng.normal(0, 0.1, n), el, -0.5)
        kw = dict(beta=3.0, gamma_start=3.0, gamma_end=0.05, num_steps=5, sweeps_per_step=2, scan_points=4,
                  scan_sweeps=2, scan_burn=1, seed=31)
        res = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_CHI_WAVE", v)
            res.append(abi.sqa_chi_run(dp, P, R, **kw))
        a, b = res
        assert np.array_equal(a["best_state"], b["best_state"]), (n, P, R, dup)
        assert bits_equal([a["best_energy"]], [b["best_energy"]]), (n, P, R, dup)
        assert bits_equal(a["energy_trace"], b["energy_trace"]), (n, P, R, dup)
        assert bits_equal(a["scan_chi_B"], b["scan_chi_B"]), (n, P, R, dup)
        pil = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_CHI_WAVE", v)
            pil.append(abi.chi_pilot(dp, P, R, 2.0, 3.0, 1e-3, 16, 5, 3, seed=43))
        assert bits_equal(pil[0]["scan_chi_B"], pil[1]["scan_chi_B"]), (n, P, R, dup)
    monkeypatch.delenv("QANNEAL_CHI_WAVE", raising=False)


def test_sqa_chi_wave_kernel_large_dense_hubo_and_sparse_pilot_bit_identical(ctx, monkeypatch):
    """The wavefront checkerboard kernel past 2048 dense sites (rows loaded on demand, 16 and 32
    float4 groups per lane), for HUBO problems (lists staged in LDS for production, read from
    global memory for the pilot scan) and for the sparse pilot scan (lists from global memory):
    bit-identical to the two-barrier kernel, and the returned energy is the oracle's FP64 energy."""
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    from tests.golden_io import consolidate_hubo
    orc = Oracle(ORACLE_SO)
    rng = np.random.default_rng(57)
    kw = dict(beta=2.5, gamma_start=3.0, gamma_end=0.05, num_steps=4, sweeps_per_step=2, scan_points=4,
              scan_sweeps=2, scan_burn=1, seed=61)
    cases = []
    for n, P, R in [(3001, 8, 2), (4096, 5, 2), (8192, 2, 1)]:
        h, J = rng.normal(0, 0.2, n), _sk(rng, n)
        cases.append((f"dense{n}", P, R, Problem.dense(h, J, 0.25), abi.DeviceProblem.dense(ctx, h, J, 0.25)))
    for n, P, R in [(60, 7, 3), (200, 8, 2)]:
        terms = _hubo_case(rng, n, 3 * n)
        hh, off, var, co, dc = consolidate_hubo(n, terms)
        cases.append((f"hubo{n}", P, R, Problem.hubo(n, terms, -0.25),
                      abi.DeviceProblem.hubo(ctx, n, hh, off, var, co, -0.25 + dc)))
    for name, P, R, po, pd in cases:
        print("chi wave case", name, P, R, flush=True)
        res = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_CHI_WAVE", v)
            res.append(abi.sqa_chi_run(pd, P, R, **kw))
        a, b = res
        assert np.array_equal(a["best_state"], b["best_state"]), name
        assert bits_equal([a["best_energy"]], [b["best_energy"]]), name
        assert bits_equal(a["energy_trace"], b["energy_trace"]), name
        assert bits_equal(a["scan_chi_B"], b["scan_chi_B"]), name
        assert a["best_energy"] == orc.energy(po, a["best_state"])[0], name
        pil = []
        for v in ("1", "0"):
            monkeypatch.setenv("QANNEAL_CHI_WAVE", v)
            pil.append(abi.chi_pilot(pd, P, R, 2.0, 3.0, 1e-3, 8, 3, 2, seed=47))
        assert bits_equal(pil[0]["scan_chi_B"], pil[1]["scan_chi_B"]), name
    monkeypatch.delenv("QANNEAL_CHI_WAVE", raising=False)


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


def _sparse_case(rng, n, density):
    ei, ej = np.triu_indices(n, 1)
    keep = rng.random(ei.size) < density
    return rng.normal(0, 0.3, n), [(int(a), int(b), float(v)) for a, b, v in
                                   zip(ei[keep], ej[keep], rng.normal(0, 1, int(keep.sum())))]


def _hubo_case(rng, n, nterms):
    """Raw HigherOrderIsing terms: distinct 3- and 4-body supports with N(0,1) coefficients plus
    N(0, 0.3) linear terms."""
    seen, terms = set(), []
    while len(terms) < nterms:
        t = tuple(sorted(rng.choice(n, int(rng.integers(3, 5)), replace=False).tolist()))
        if t not in seen:
            seen.add(t)
            terms.append((t, float(rng.normal())))
    return terms + [((i,), float(rng.normal(0, 0.3))) for i in range(n)]


def test_sqa_fp32_sparse_hubo(ctx):
    """Sparse and higher-order couplings through the FP32 two-barrier sweep (slice, worldline and
    cluster moves): every returned best energy is the oracle's FP64 energy of the returned state,
    and the best energies of independent reads follow the oracle SQA's distribution (two-sample
    KS, p > 0.01)."""
    from scipy.stats import ks_2samp
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    from tests.golden_io import consolidate_hubo
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    rng = np.random.default_rng(23)
    h, edges = _sparse_case(rng, 160, 0.04)
    sp_o, sp_d = Problem.sparse(h, edges, 0.5), abi.DeviceProblem.sparse(ctx, h, edges, 0.5)
    terms = _hubo_case(rng, 50, 150)
    hb_o = Problem.hubo(50, terms, -0.25)
    hh, off, var, co, dc = consolidate_hubo(50, terms)
    hb_d = abi.DeviceProblem.hubo(ctx, 50, hh, off, var, co, -0.25 + dc)
    for name, po, pd in (("sparse", sp_o, sp_d), ("hubo", hb_o, hb_d)):
        for P, R, wl, cl in ((8, 3, 0, 0), (6, 2, 1, 1), (1, 4, 0, 1)):
            print("exact-energy case", name, P, R, wl, cl, flush=True)
            r = abi.sqa_run(pd, np.full(12, 2.0), np.geomspace(2.0, 0.05, 12), P, R, 1, wl, cl, abi.MODE_FP32, 7)
            assert r["best_energy"] == orc.energy(po, r["best_state"])[0], (name, P, wl, cl)
    # distributions: small instances, 128 independent reads per path
    h, edges = _sparse_case(rng, 24, 0.3)
    terms = _hubo_case(rng, 16, 30)
    hh, off, var, co, dc = consolidate_hubo(16, terms)
    cases = (("sparse", Problem.sparse(h, edges), abi.DeviceProblem.sparse(ctx, h, edges)),
             ("hubo", Problem.hubo(16, terms), abi.DeviceProblem.hubo(ctx, 16, hh, off, var, co, dc)))
    steps, P = 40, 8
    betas, gammas = np.full(steps, 1.5), np.linspace(2.0, 0.05, steps)
    for name, po, pd in cases:
        gpu = np.array([abi.sqa_run(pd, betas, gammas, P, 1, 1, 1, 1, abi.MODE_FP32, 300 + k)["best_energy"]
                        for k in range(128)])
        cpu = np.array([orc.sqa_run(po, betas, gammas, P, 1, 1, 1, 1, 0, 900 + k)["best_energy"]
                        for k in range(128)])
        p = ks_2samp(gpu, cpu).pvalue
        print("KS", name, p, "gpu mean", gpu.mean(), "cpu mean", cpu.mean(), flush=True)
        assert p > 0.01, name


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


def test_sqapt_validation_golden(ctx):
    from qlatanneal_amd import abi
    for c in load("sqapt.json")["cases"]:
        print("sqapt case", c["seed"], flush=True)
        dp = device_problem(ctx, c["problem"])
        betas, gammas = dec(c["betas"]), dec(c["gammas"])
        R = len(betas)
        P = max(c["trotter_slices"], c["continuous_time_slices"])
        g = MT64(c["seed"])
        init = g.spins(R * P * dp.n)
        seeds = g.raw(R)
        swap_u = g.uniforms(c["steps"] * (R - 1))
        r = abi.sqapt_run(dp, betas, gammas, P, c["sweeps_per_step"], c["worldline_sweeps"], c["steps"],
                          c["swap_interval"], c["cluster_sweeps"], 1, abi.MODE_FP64_VALIDATION, 0, init, seeds, swap_u)
        exp = out(c["out"])
        assert bits_equal(r["final_states"][0], exp["final_states"].reshape(R, -1))
        assert bits_equal(r["final_energies"][0], exp["final_energies"])
        assert bits_equal(r["best_state"][0], exp["best_state"])
        assert bits_equal(r["best_energy"], exp["best_energy"])
        assert bits_equal(r["average_energy_trace"][0], exp["average_energy_trace"])
        assert bits_equal(r["swap_acceptance_trace"][0], exp["swap_acceptance_trace"])


def test_sqapt_fp32_exact_energies_and_distribution(ctx):
    from scipy.stats import ks_2samp
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(23)
    n, P, steps = 32, 8, 25
    J = _sk(rng, n)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    betas, gammas = np.geomspace(0.3, 3.0, 6), np.geomspace(3.0, 0.05, 6)
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    prob = Problem.dense(np.zeros(n), J)
    # one launch with 256 independent ladders vs 256 seeded reference runs
    r = abi.sqapt_run(dp, betas, gammas, P, 1, 1, steps, 1, 0, ladders=256, mode=abi.MODE_FP32, seed=3)
    for l in range(4):
        assert r["best_energy"][l] == orc.energy(prob, r["best_state"][l])[0]
        for q in range(len(betas)):
            assert r["final_energies"][l, q] == orc.energy(prob, r["final_states"][l, q])[0]
    cpu = np.array([orc.sqapt_run(prob, betas, gammas, P, 1, 1, steps, 1, 0, 0, 900 + k)["best_energy"] for k in range(256)])
    p = ks_2samp(r["best_energy"], cpu).pvalue
    print("PT KS p", p, r["best_energy"].mean(), cpu.mean(), "swap acc", r["swap_acceptance_trace"].mean(), flush=True)
    assert p > 0.01


def test_sqa_chi_validation_golden(ctx):
    from qlatanneal_amd import abi
    for c in load("sqa_chi.json")["cases"]:
        print("chi case", c["seed"], c["trotter_slices"], flush=True)
        dp = device_problem(ctx, c["problem"])
        P, R, SP = c["trotter_slices"], c["replicas"], c["scan_points"]
        g = MT64(c["seed"])
        init, seeds = [], []
        for _ in range(SP + 1):
            init.append(g.spins(R * P * dp.n))
            seeds.append(g.raw(R * P))
        r = abi.sqa_chi_run(dp, P, R, c["beta"], c["gamma_start"], c["gamma_end"], c["num_steps"],
                            c["sweeps_per_step"], SP, c["scan_sweeps"], c["scan_burn"],
                            mode=abi.MODE_FP64_VALIDATION, init_spins=np.concatenate(init),
                            cell_seeds=np.concatenate(seeds))
        exp = out(c["out"])
        for k in exp:
            assert bits_equal(r[k], exp[k]), (c["seed"], k)


def test_sqa_chi_fp32(ctx):
    from scipy.stats import ks_2samp
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    # dense ferromagnetic chain: exact ground state -(n-1) (tests/test_sqa_chi_annealer.cpp:241-255)
    n = 12
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = J[i + 1, i] = -1.0
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    r = abi.sqa_chi_run(dp, 8, 4, 4.0, 3.0, 0.05, 100, 10, 10, 20, 6, seed=7)
    assert r["best_energy"] == -(n - 1)
    assert np.all(np.diff(r["gamma_schedule"]) <= 1e-12)
    assert abs(r["gamma_schedule"][0] - 3.0) < 1e-9 and abs(r["gamma_schedule"][-1] - 0.05) < 1e-6
    assert np.all(r["scan_chi_B"] >= 0) and r["scan_chi_B"].max() > r["scan_chi_B"].min()
    # SK instance: residual energies vs the oracle's run_chi over 256 seeds
    rng = np.random.default_rng(31)
    n = 24
    J = _sk(rng, n)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    prob = Problem.dense(np.zeros(n), J)
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    kw = dict(beta=2.0, gamma_start=3.0, gamma_end=0.05, num_steps=20, sweeps_per_step=2, scan_points=8,
              scan_sweeps=6, scan_burn=2)
    runs = [abi.sqa_chi_run(dp, 4, 2, seed=100 + k, **kw) for k in range(256)]
    gpu = np.array([r["best_energy"] for r in runs])
    cpu = np.array([orc.sqa_chi_run(prob, 4, 2, seed=700 + k, **kw)["best_energy"] for k in range(256)])
    # every returned sample's energy is the FP64 energy of its state
    states = np.stack([r["best_state"] for r in runs])
    assert np.all(np.isfinite(gpu))
    assert np.array_equal(gpu, dp.energy(states))
    p = ks_2samp(gpu, cpu).pvalue
    print("chi KS p", p, gpu.mean(), cpu.mean(), flush=True)
    assert p > 0.01


def test_sqapt_fp32_c3_shape_with_refresh(ctx):
    """SQA-PT at the C3 shape — dense SK N=2048, P=64, 32 rungs x 4 ladders, 10 slice sweeps per
    step, field refresh on the matrix cores — for a short schedule: every ladder's best and final
    energies are the exact FP64 energies of the returned states, the drift after a refresh stays at
    the split bound, and swap acceptances are probabilities."""
    import time
    from qlatanneal_amd import abi
    rng = np.random.default_rng(31)
    n, P, rungs, L, steps = 2048, 64, 32, 4, 6
    J = _sk(rng, n)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    betas = np.geomspace(0.5, 4.0, rungs)
    gammas = np.geomspace(3.0, 0.05, rungs)
    t = time.time()
    r = abi.sqapt_run(dp, betas, gammas, P, 10, 0, steps, 1, 0, ladders=L, mode=abi.MODE_FP32, seed=3,
                      refresh_interval=20)
    wall = time.time() - t
    attempts = steps * 10 * rungs * L * P * n
    print(f"C3 shape: {steps} steps, kernel {r['kernel_ms']:.0f} ms, wall {wall:.1f} s, "
          f"{attempts / (r['kernel_ms'] / 1e3):.3e} attempts/s, drift {r['field_drift']}", flush=True)
    for l in range(L):
        assert r["best_energy"][l] == dp.energy(r["best_state"][l])[0]
        assert np.array_equal(r["final_energies"][l], dp.energy(r["final_states"][l]))
    assert np.all((r["swap_acceptance_trace"] >= 0) & (r["swap_acceptance_trace"] <= 1))
    before, after = r["field_drift"]
    assert np.isfinite(before) and after <= 5e-4


def _maxcut3(rng, n):
    """Unit-weight 3-regular-style graph: edges from random perfect matchings, 3n/2 edges."""
    edges = set()
    while len(edges) < 3 * n // 2:
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            u, v = sorted((int(perm[i]), int(perm[i + 1])))
            if len(edges) < 3 * n // 2:
                edges.add((u, v))
    return [(u, v, 1.0) for u, v in sorted(edges)]


def test_sqa_fp32_full_c2_shape_exact_and_distribution(ctx):
    """The headline SQA shape — dense SK N=512, P=32, 256 replicas, beta=32, linear Gamma — for a
    20-sweep schedule: the returned best energy is the FP64 energy of the returned state, the
    trace is finite and the acceptance is a rate. Same P on a smaller N (64): best energies over
    256 independent reads agree with the oracle's (two-sample KS)."""
    from scipy.stats import ks_2samp
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(512)
    n, P, R, steps = 512, 32, 256, 20
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(rng, n))
    betas, gammas = np.full(steps, 32.0), np.linspace(3.0, 1e-3, steps)
    for seed in (1, 2):
        r = abi.sqa_run(dp, betas, gammas, P, R, 1, 0, 0, abi.MODE_FP32, seed)
        print("C2 shape seed", seed, "best", r["best_energy"], "kernel ms", r["kernel_ms"], flush=True)
        assert r["best_energy"] == dp.energy(r["best_state"])[0]
        assert np.all(np.isfinite(r["energy_trace"])) and r["energy_trace"][-1] < r["energy_trace"][0]
        assert 0.0 < r["accept_rate"] < 1.0
    n = 64
    J = _sk(rng, n)
    dps = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    gpu = np.array([abi.sqa_run(dps, betas, gammas, P, 1, 1, 0, 0, abi.MODE_FP32, 100 + k)["best_energy"]
                    for k in range(256)])
    cpu = np.array([orc.sqa_run(Problem.dense(np.zeros(n), J), betas, gammas, P, 1, 1, 0, 0, 0, 900 + k)["best_energy"]
                    for k in range(256)])
    p = ks_2samp(gpu, cpu).pvalue
    print("C2 P=32 N=64 KS p", p, gpu.mean(), cpu.mean(), flush=True)
    assert p > 0.01


def test_sqa_chi_fp32_full_c5_shapes_exact_and_distribution(ctx):
    """sqa_chi at the C5 shapes — dense SK N=2048 and 3-regular unit MAX-CUT N=4096 (sparse, the
    adjacency staged in LDS), P=32, R=4, 64 scan points over Gamma 3.0 -> 1e-3 — with a short pilot
    and production: the best energy is the FP64 energy of the best state, the Gamma schedule runs
    monotonically from gamma_start to gamma_end, the scan is nonnegative. Same P, R and scan size
    on small instances of both families: best energies over 128 runs agree with the oracle (KS)."""
    from scipy.stats import ks_2samp
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    rng = np.random.default_rng(2048)
    P, R = 32, 4
    kw = dict(num_steps=10, sweeps_per_step=2, scan_points=64, scan_sweeps=2, scan_burn=1)
    sk_J = _sk(rng, 2048)
    cases = [("sk2048", abi.DeviceProblem.dense(ctx, np.zeros(2048), sk_J)),
             ("maxcut4096", abi.DeviceProblem.sparse(ctx, np.zeros(4096), _maxcut3(rng, 4096)))]
    for name, dp in cases:
        r = abi.sqa_chi_run(dp, P, R, 4.0, 3.0, 1e-3, seed=5, **kw)
        print("C5 shape", name, "best", r["best_energy"], "kernel ms", r["kernel_ms"], flush=True)
        assert r["best_energy"] == dp.energy(r["best_state"])[0], name
        g = r["gamma_schedule"]
        assert np.all(np.diff(g) <= 1e-12) and abs(g[0] - 3.0) < 1e-9 and abs(g[-1] - 1e-3) < 1e-9, name
        assert np.all(r["scan_chi_B"] >= 0) and len(r["scan_chi_B"]) == 64, name
    orc = Oracle(ORACLE_SO)
    orc.set_threads(8)
    for name, n in (("sk", 32), ("maxcut", 64)):
        if name == "sk":
            J = _sk(rng, n)
            dp, prob = abi.DeviceProblem.dense(ctx, np.zeros(n), J), Problem.dense(np.zeros(n), J)
        else:
            e = _maxcut3(rng, n)
            dp, prob = abi.DeviceProblem.sparse(ctx, np.zeros(n), e), Problem.sparse(np.zeros(n), e)
        gpu = np.array([abi.sqa_chi_run(dp, P, R, 4.0, 3.0, 1e-3, seed=300 + k, **kw)["best_energy"]
                        for k in range(128)])
        cpu = np.array([orc.sqa_chi_run(prob, P, R, 4.0, 3.0, 1e-3, seed=800 + k, **kw)["best_energy"]
                        for k in range(128)])
        p = ks_2samp(gpu, cpu).pvalue
        print("C5", name, n, "KS p", p, gpu.mean(), cpu.mean(), flush=True)
        assert p > 0.01, name
