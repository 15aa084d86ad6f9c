"""Annealed SQA-PT ladders (qa_sqapt_params.step_betas / step_gammas) and the schedule-only
sqa_chi entry point (qa_chi_schedule).

The reference's SQA parallel tempering keeps one ladder for a whole run
(sqa_parallel_tempering.cpp:97-200); a per-step ladder whose rows all equal that ladder must
therefore reproduce the fixed-ladder run bit for bit, on every sweep kernel the launch can pick.
qa_chi_schedule is run_chi's first half (sqa_chi_annealer.cpp:146-330) and must return the
schedules qa_sqa_chi_run anneals on, bit for bit, for each source of the chi_B profile.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "qlatanneal_amd", "data", "chi_surrogate.npz")


@pytest.fixture(scope="module")
def ctx():
    from qlatanneal_amd.abi import Context
    return Context(0)


def _sk(n, seed):
    rng = np.random.default_rng(seed)
    J = np.triu(rng.normal(0.0, 1.0 / np.sqrt(n), (n, n)), 1)
    return J + J.T


KEYS = ("final_states", "final_energies", "best_state", "best_energy", "average_energy_trace",
        "swap_acceptance_trace")


# (n, P, rungs, ladders, steps, sweeps): two-barrier / pipelined kernels at small n, the column
# loop at n > 1024 with P = 8, the LDS-state look-ahead kernel at n = 2048, P = 32
@pytest.mark.parametrize("n,P,R,L,steps,spb", [(256, 16, 6, 3, 7, 5), (1100, 8, 4, 2, 4, 3), (2048, 32, 4, 2, 3, 2)])
def test_constant_step_ladder_equals_fixed_ladder(ctx, n, P, R, L, steps, spb):
    from qlatanneal_amd import abi
    print(f"case n={n} P={P} R={R} L={L} steps={steps}", flush=True)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(n, n))
    b, g = np.geomspace(0.5, 4.0, R), np.geomspace(2.0, 0.05, R)
    fixed = abi.sqapt_run(dp, b, g, P, spb, 0, steps, 1, 0, ladders=L, seed=11)
    rows = abi.sqapt_run(dp, b, g, P, spb, 0, steps, 1, 0, ladders=L, seed=11, step_betas=np.tile(b, (steps, 1)),
                         step_gammas=np.tile(g, (steps, 1)))
    for k in KEYS:
        assert np.array_equal(fixed[k], rows[k]), k


def test_step_ladder_rows_drive_sweeps_and_exchanges(ctx):
    """Rows 0-3: every rung at one hot (beta, Gamma), so each exchange has a zero action difference
    and is accepted; rows 4-7: a wide cold ladder. A launch that read row 0 for every step would
    stay hot and accept every exchange."""
    from qlatanneal_amd import abi
    n, P, R, L, steps = 256, 16, 6, 4, 8
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(n, 3))
    sb = np.empty((steps, R))
    sg = np.empty((steps, R))
    sb[:4], sg[:4] = 0.05, 1.0
    sb[4:], sg[4:] = np.geomspace(2.0, 8.0, R), np.geomspace(0.5, 0.01, R)
    r = abi.sqapt_run(dp, sb[-1], sg[-1], P, 4, 0, steps, 1, 0, ladders=L, seed=5, step_betas=sb, step_gammas=sg)
    sw, av = r["swap_acceptance_trace"], r["average_energy_trace"]
    print("swap", sw.mean(axis=0), "avg", av.mean(axis=0), flush=True)
    assert np.all(sw[:, :4] == 1.0)
    assert sw[:, 4:].mean() < 0.9
    assert av[:, 7].max() < av[:, 3].min() - 0.2 * n
    for l in range(L):
        assert r["best_energy"][l] == dp.energy(r["best_state"][l])[0]


def test_run_ladders_class_step_ladder(ctx):
    import qlatanneal_amd as qa
    n, P, R, steps = 128, 8, 4, 5
    J = _sk(n, 9)
    h = qa.DenseIsing(np.zeros(n), J)
    b, g = np.geomspace(0.5, 4.0, R).tolist(), np.geomspace(2.0, 0.05, R).tolist()
    a = qa.SQAParallelTemperingAnnealer(h, b, g, P, backend="hip")
    a.set_seed(4)
    fixed = a.run_ladders(3, 4, 0, steps, 1, 0)
    a.set_seed(4)
    rows = a.run_ladders(3, 4, 0, steps, 1, 0, step_betas=[b] * steps, step_gammas=[g] * steps)
    for x, y in zip(fixed, rows):
        assert x.best_energy == y.best_energy and list(x.best_state.spins) == list(y.best_state.spins)
        assert list(x.average_energy_trace) == list(y.average_energy_trace)
    with pytest.raises(ValueError):
        a.run_ladders(3, 4, 0, steps, 1, 0, step_betas=[b] * (steps - 1), step_gammas=[g] * (steps - 1))
    with pytest.raises(ValueError):
        a.run_ladders(3, 4, 0, steps, 1, 0, step_betas=[b] * steps)


def _schedule_keys():
    return ("beta_schedule", "gamma_schedule", "s_schedule", "scan_s", "scan_gamma", "scan_chi_B", "s_star",
            "gamma_star", "j_perp_star", "chi_floor", "driver_A0")


def test_chi_schedule_equals_sqa_chi_run(ctx):
    from qlatanneal_amd import abi
    n = 64
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(n, 21))
    kw = dict(scan_points=8, scan_sweeps=6, scan_burn=2, driver_A0=3.0, seed=77, beta_ramp_fraction=0.3,
              beta_ramp_start=0.1)
    for extra in ({}, dict(chi_override=np.linspace(0.2, 1.0, 8) ** 2)):
        sch = abi.chi_schedule(dp, 8, 2, 4.0, 3.0, 1e-3, num_steps=20, **kw, **extra)
        run = abi.sqa_chi_run(dp, 8, 2, 4.0, 3.0, 1e-3, num_steps=20, sweeps_per_step=2, **kw, **extra)
        for k in _schedule_keys():
            assert np.array_equal(sch[k], run[k]), k
        assert sch["surrogate_used"] == run["surrogate_used"] is False


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained chi_B surrogate in the package")
def test_chi_schedule_with_trained_surrogate(ctx):
    from qlatanneal_amd import abi
    from qlatanneal_amd import surrogate as S
    sur = S.ChiSurrogate.load(ctx, MODEL)
    cfg = sur.cfg
    n = 512
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(n, 22))
    kw = dict(scan_points=cfg.scan_points, scan_sweeps=cfg.scan_sweeps, scan_burn=cfg.scan_burn,
              driver_A0=cfg.driver_A0, beta_ramp_fraction=0.0, surrogate=sur.mlp, seed=3)
    sch = abi.chi_schedule(dp, cfg.trotter_slices, cfg.replicas, 16.0, cfg.gamma_start, cfg.gamma_end,
                           num_steps=40, **kw)
    run = abi.sqa_chi_run(dp, cfg.trotter_slices, cfg.replicas, 16.0, cfg.gamma_start, cfg.gamma_end,
                          num_steps=40, sweeps_per_step=1, **kw)
    assert sch["surrogate_used"] and run["surrogate_used"]
    for k in _schedule_keys():
        assert np.array_equal(sch[k], run[k]), k
    assert np.all(sch["beta_schedule"] == 16.0)
    assert sch["gamma_schedule"][0] == cfg.gamma_start and np.all(np.diff(sch["gamma_schedule"]) <= 0.0)


def test_annealed_ladder_exact_ground_states(ctx):
    """Ferromagnetic chain (tests/test_sqa_chi_annealer.cpp:241-271's instance): every ladder of an
    annealed SQA-PT launch on the pilot schedule ends in the ground state, -(n - 1)."""
    from qlatanneal_amd import abi
    from qlatanneal_amd.surrogate import chi_annealed_ladder
    n = 12
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = J[i + 1, i] = -1.0
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    sch = abi.chi_schedule(dp, 8, 4, 4.0, 3.0, 0.05, num_steps=100, scan_points=10, scan_sweeps=20, scan_burn=6,
                           seed=7)
    sb, sg = chi_annealed_ladder(sch["beta_schedule"], sch["gamma_schedule"], 4, 0.5)
    r = abi.sqapt_run(dp, sb[-1], sg[-1], 8, 10, 0, 100, 1, 0, ladders=6, seed=8, step_betas=sb, step_gammas=sg)
    assert np.all(r["best_energy"] == -(n - 1))


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained chi_B surrogate in the package")
def test_annealed_ladder_sk2048_within_one_percent_of_pilot_sqa_chi(ctx):
    """The target configuration, bounded: dense SK N=2048, the surrogate's Gamma schedule at beta=16,
    4 rungs (beta 8 -> 16), 100 x 10 sweeps, FP32 fields refreshed every 100 sweeps, 8 ladders in
    one launch. Every returned energy is the FP64 energy of its state, and the ladders' mean is
    within 1 % of the pilot-scheduled sqa_chi's (P=32, R=4, 100 x 10 sweeps)."""
    from qlatanneal_amd import abi
    from qlatanneal_amd import surrogate as S
    sur = S.ChiSurrogate.load(ctx, MODEL)
    cfg = sur.cfg
    n = 2048
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(n, 20261210))
    chi_kw = dict(scan_points=cfg.scan_points, scan_sweeps=cfg.scan_sweeps, scan_burn=cfg.scan_burn,
                  driver_A0=cfg.driver_A0)
    pilot = abi.sqa_chi_run(dp, cfg.trotter_slices, cfg.replicas, 16.0, cfg.gamma_start, cfg.gamma_end,
                            num_steps=100, sweeps_per_step=10, seed=31, **chi_kw)
    sch = abi.chi_schedule(dp, cfg.trotter_slices, cfg.replicas, 16.0, cfg.gamma_start, cfg.gamma_end,
                           num_steps=100, beta_ramp_fraction=0.0, surrogate=sur.mlp, seed=5, **chi_kw)
    assert sch["surrogate_used"]
    sb, sg = S.chi_annealed_ladder(sch["beta_schedule"], sch["gamma_schedule"], 4, 0.5)
    r = abi.sqapt_run(dp, sb[-1], sg[-1], cfg.trotter_slices, 10, 0, 100, 1, 0, ladders=8, seed=6,
                      step_betas=sb, step_gammas=sg, refresh_interval=100)
    e = r["best_energy"]
    print("pilot sqa_chi", pilot["best_energy"] / n, "ladders", e / n, "drift", r["field_drift"], flush=True)
    for l in range(8):
        assert e[l] == dp.energy(r["best_state"][l])[0]
    assert e.mean() <= pilot["best_energy"] + 0.01 * abs(pilot["best_energy"])
    assert np.all(np.isfinite(r["field_drift"])) and r["field_drift"][1] <= 5e-4


def test_step_ladder_argument_errors(ctx):
    """A per-step ladder is an FP32 input with positive values and one row per step; the schedule
    entry point is FP32-only."""
    from qlatanneal_amd import abi
    n, P, R, steps = 32, 4, 3, 4
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), _sk(n, 1))
    b, g = np.geomspace(0.5, 2.0, R), np.geomspace(1.0, 0.1, R)
    sb, sg = np.tile(b, (steps, 1)), np.tile(g, (steps, 1))
    bad = sb.copy()
    bad[2, 1] = 0.0
    with pytest.raises(abi.QaInvalidArgument):
        abi.sqapt_run(dp, b, g, P, 1, 0, steps, 1, 0, seed=1, step_betas=bad, step_gammas=sg)
    bad = sg.copy()
    bad[1, 0] = np.nan
    with pytest.raises(abi.QaInvalidArgument):
        abi.sqapt_run(dp, b, g, P, 1, 0, steps, 1, 0, seed=1, step_betas=sb, step_gammas=bad)
    with pytest.raises(ValueError):
        abi.sqapt_run(dp, b, g, P, 1, 0, steps, 1, 0, seed=1, step_betas=sb[:-1], step_gammas=sg[:-1])
    with pytest.raises(abi.QaError) as e:
        abi.sqapt_run(dp, b, g, P, 1, 0, steps, 1, 0, mode=abi.MODE_FP64_VALIDATION, seed=1, step_betas=sb,
                      step_gammas=sg, init_spins=np.ones(R * P * n, np.int8), rung_seeds=np.arange(R, dtype=np.uint64),
                      swap_uniforms=np.full(steps * (R - 1), 0.5))
    assert e.value.status == abi.QA_EUNSUPPORTED
