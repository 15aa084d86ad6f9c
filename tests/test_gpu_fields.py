"""Matrix-core local fields (qa_local_fields) against the oracle's FP64 fields.

Dense symmetric couplings run as an FP32-accumulated bf16 GEMM with J split into hi + lo parts.
Each field is checked against the oracle (the reference's DenseIsing::compute_local_fields order
in FP64) within a bound built from the actual split residuals: sum_k |J_jk - hi_jk - lo_jk| (the
split), sum_k |J_jk - fl32(J_jk)| (FP32 couplings) and an FP32 accumulation term. Couplings with
at most 16 significant bits (QUBO-derived half-integers) must come out exact. Sparse and HUBO
models return the FP64 reference-order fields rounded once to FP32, bit for bit.
"""
import numpy as np
import pytest

from tests.golden_io import load, problem, device_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from qlatanneal_amd.abi import Context
    return Context(0)


def _bf16_rne(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _sk(rng, n):
    J = rng.normal(0, 1 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def test_local_fields_dense_within_split_bound(ctx):
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    orc = Oracle(ORACLE_SO)
    rng = np.random.default_rng(7)
    for n, count in [(64, 64), (100, 70), (257, 130), (512, 33), (31, 5)]:
        print("dense fields case", n, count, flush=True)
        J = _sk(rng, n)
        h = rng.normal(0, 0.3, n)
        dp = abi.DeviceProblem.dense(ctx, h, J, 0.0)
        S = rng.choice(np.array([-1, 1], np.int8), size=(count, n))
        F = dp.local_fields(S).astype(np.float64)
        prob = Problem.dense(h, J)
        ref = np.stack([orc.local_fields(prob, S[m]) for m in range(count)])
        J32 = J.astype(np.float32)
        np.fill_diagonal(J32, 0.0)
        hi = _bf16_rne(J32)
        lo = _bf16_rne(J32 - hi)
        res = J32.astype(np.float64) - hi - lo
        r32 = J - J32.astype(np.float64)
        np.fill_diagonal(r32, 0.0)
        absS = np.abs(S.astype(np.float64))
        bound = (absS @ (np.abs(res) + np.abs(r32)).T + 2.0 ** -22 * (n + 2) * (np.abs(h) + absS @ np.abs(J).T)
                 + np.abs(h - h.astype(np.float32)))
        err = np.abs(F - ref)
        print("  max err", err.max(), "max bound", bound.max(), flush=True)
        assert np.all(err <= bound)
        # a transposed or shifted tile would be off by O(|J| sqrt(n)) = O(1), far above the bound
        assert bound.max() < 0.01


def test_local_fields_exact_for_half_integer_couplings(ctx):
    from qlatanneal_amd import abi
    rng = np.random.default_rng(11)
    for n, count in [(96, 40), (200, 65)]:
        Q = rng.integers(-100, 101, (n, n)).astype(np.float64)
        J = np.triu(Q, 1) / 2.0
        J = J + J.T
        h = rng.integers(-400, 401, n) / 4.0
        dp = abi.DeviceProblem.dense(ctx, h, J, 0.0)
        S = rng.choice(np.array([-1, 1], np.int8), size=(count, n))
        F = dp.local_fields(S).astype(np.float64)
        exact = h[None, :] + S.astype(np.float64) @ J.T
        assert np.array_equal(F, exact)


def test_local_fields_sparse_hubo_reference_order(ctx):
    from oracle.orc import Oracle, ORACLE_SO, DENSE
    orc = Oracle(ORACLE_SO)
    rng = np.random.default_rng(5)
    for c in load("hamiltonian.json")["cases"]:
        p = problem(c["problem"])
        if p.kind == DENSE:
            continue
        dp = device_problem(ctx, c["problem"])
        S = rng.choice(np.array([-1, 1], np.int8), size=(6, p.n))
        F = dp.local_fields(S)
        for m in range(6):
            want = orc.local_fields(p, S[m]).astype(np.float32)
            assert np.array_equal(F[m], want)


def test_refresh_is_identity_for_exact_fields(ctx):
    """With half-integer couplings every FP32 field is exact, so a matrix-core refresh reproduces
    the running fields bit for bit: runs with and without refreshes must be identical, and the
    drift check must read 0 before and after each refresh."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(13)
    n, P, R = 80, 8, 6
    Q = rng.integers(-20, 21, (n, n)).astype(np.float64)
    J = np.triu(Q, 1) / 2.0
    J = J + J.T
    h = rng.integers(-40, 41, n) / 4.0
    dp = abi.DeviceProblem.dense(ctx, h, J, 0.0)
    betas, gammas = np.full(10, 0.5), np.geomspace(3.0, 0.05, 10)
    a = abi.sqa_run(dp, betas, gammas, P, R, 2, 0, 0, abi.MODE_FP32, 21)
    b = abi.sqa_run(dp, betas, gammas, P, R, 2, 0, 0, abi.MODE_FP32, 21, refresh_interval=4)
    assert np.isnan(a["field_drift"]).all()
    assert np.array_equal(b["field_drift"], [0.0, 0.0])
    assert a["accept_rate"] == b["accept_rate"]
    assert np.array_equal(a["best_state"], b["best_state"])
    assert np.array_equal(a["energy_trace"], b["energy_trace"])
    lad = dict(ladders=4, mode=abi.MODE_FP32, seed=8)
    pa = abi.sqapt_run(dp, np.geomspace(0.3, 3.0, 4), np.geomspace(3.0, 0.05, 4), P, 2, 0, 8, 1, 0, **lad)
    pb = abi.sqapt_run(dp, np.geomspace(0.3, 3.0, 4), np.geomspace(3.0, 0.05, 4), P, 2, 0, 8, 1, 0,
                       refresh_interval=3, **lad)
    assert np.array_equal(pb["field_drift"], [0.0, 0.0])
    for key in ("final_states", "best_state", "final_energies", "average_energy_trace", "swap_acceptance_trace"):
        assert np.array_equal(pa[key], pb[key]), key


def test_refresh_bounds_drift_and_keeps_samples_exact(ctx):
    """Gaussian couplings: the running FP32 fields drift from the FP64 fields; after each refresh
    the error is back to the split bound, and every returned sample still carries its exact FP64
    energy."""
    from qlatanneal_amd import abi
    from oracle.orc import Oracle, Problem, ORACLE_SO
    orc = Oracle(ORACLE_SO)
    rng = np.random.default_rng(17)
    n, P, R = 256, 16, 8
    J = _sk(rng, n)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J, 0.0)
    betas, gammas = np.full(40, 2.0), np.linspace(2.0, 0.05, 40)
    r = abi.sqa_run(dp, betas, gammas, P, R, 1, 0, 0, abi.MODE_FP32, 3, refresh_interval=10)
    before, after = r["field_drift"]
    print("drift before/after refresh", before, after, flush=True)
    assert np.isfinite(before) and np.isfinite(after)
    assert after <= 2e-4
    assert r["best_energy"] == orc.energy(Problem.dense(np.zeros(n), J), r["best_state"])[0]
    p = abi.sqapt_run(dp, np.geomspace(0.5, 3.0, 4), np.geomspace(2.0, 0.05, 4), P, 2, 0, 12, 1, 0,
                      ladders=4, mode=abi.MODE_FP32, seed=4, refresh_interval=5)
    assert np.isfinite(p["field_drift"]).all() and p["field_drift"][1] <= 2e-4
    prob = Problem.dense(np.zeros(n), J)
    for l in range(4):
        assert p["best_energy"][l] == orc.energy(prob, p["best_state"][l])[0]


@pytest.mark.parametrize("n,P,R", [(300, 3, 2), (2048, 64, 2)])
def test_drift_fields_kernel_equals_one_state_per_workgroup(ctx, n, P, R, monkeypatch):
    """The drift check's FP64 fields (fields64_sym_kernel: 64 elements x 8 states per workgroup,
    J read down columns) are the same additions in the same order as the one-state-per-workgroup
    kernel, so the drift values of two otherwise identical runs are bit-equal. Nonzero h and a
    nonzero diagonal (kept in the FP64 couplings, removed from every field) exercise the first and
    last terms; n = 300 leaves a partial element block and 6 sampled chains a partial state block."""
    from qlatanneal_amd import abi
    rng = np.random.default_rng(n)
    J = _sk(rng, n) + np.diag(rng.normal(0, 1, n))
    h = rng.normal(0, 0.3, n)
    dp = abi.DeviceProblem.dense(ctx, h, J, 0.0)
    betas, gammas = np.full(6, 1.5), np.linspace(2.0, 0.1, 6)
    out = {}
    for mode in ("reference", "tiled"):
        if mode == "reference":
            monkeypatch.setenv("QANNEAL_DRIFT_FIELDS", "reference")
        else:
            monkeypatch.delenv("QANNEAL_DRIFT_FIELDS", raising=False)
        out[mode] = abi.sqa_run(dp, betas, gammas, P, R, 2, 0, 0, abi.MODE_FP32, 9, refresh_interval=4)
    a, b = out["reference"], out["tiled"]
    print("drift", a["field_drift"], b["field_drift"], flush=True)
    assert np.all(np.isfinite(a["field_drift"])) and a["field_drift"][1] <= 2e-4
    assert np.array_equal(a["field_drift"], b["field_drift"])
    assert np.array_equal(a["best_state"], b["best_state"]) and a["best_energy"] == b["best_energy"]
