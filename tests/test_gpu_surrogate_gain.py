"""What the trained chi_B surrogate gains over the method without it.

run_chi measures chi_B with a pilot scan before production (sqa_chi_annealer.cpp:232-281); the
surrogate (qlatanneal_amd/data/chi_surrogate.npz, trained on GPU pilot scans by the library's own
kernels) predicts that profile instead. On a fixed held-out set spanning the C5 families and sizes
-- SK N 256-2048 and 3-regular MAX-CUT N 512-4096, beta 2-16 -- every instance runs run_chi three
ways with the same production seed: pilot-scheduled, surrogate-scheduled, and flat (no chi_B
information: a schedule linear in s).

Pass criteria: the surrogate is used (no pilot fallback) on every instance; its mean best-energy
gap to the pilot-scheduled run is at most 1 % (gap = (E_surrogate - E_pilot) / |E_pilot|); and its
device time is at most 90 % of the pilot-scheduled runs' (the pilot, the removed cost, is 12-20 %
of run_chi at these sizes with the wavefront pilot kernel). The flat
schedule's gap is reported beside it. Every returned energy is the FP64 energy of its state.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qlatanneal_amd", "data",
                     "chi_surrogate.npz")

# (family, n, beta, instance seed): seeds outside the training draws (numpy integers in [1, 2^62))
HELD_OUT = [("sk", 256, 2.0, 11), ("sk", 1024, 8.0, 12), ("sk", 2048, 4.0, 13), ("sk", 2048, 16.0, 14),
            ("maxcut3", 512, 4.0, 15), ("maxcut3", 2048, 2.0, 16), ("maxcut3", 4096, 4.0, 17),
            ("maxcut3", 4096, 8.0, 18)]


def test_surrogate_schedule_matches_pilot_and_saves_the_pilot():
    from qlatanneal_amd import abi
    from qlatanneal_amd import surrogate as S
    assert os.path.exists(MODEL), "the trained surrogate model file is part of the package"
    ctx = abi.Context(0)
    sur = S.ChiSurrogate.load(ctx, MODEL)
    cfg = sur.cfg
    kw = dict(num_steps=100, sweeps_per_step=10, scan_points=cfg.scan_points, scan_sweeps=cfg.scan_sweeps,
              scan_burn=cfg.scan_burn, driver_A0=cfg.driver_A0)
    rows = []
    for fam, n, beta, seed in HELD_OUT:
        prob = S.build(ctx, S.InstanceSpec(fam, n, beta, seed))
        args = (prob, cfg.trotter_slices, cfg.replicas, beta, cfg.gamma_start, cfg.gamma_end)
        pil = abi.sqa_chi_run(*args, seed=1000 + seed, **kw)
        sr = abi.sqa_chi_run(*args, surrogate=sur.mlp, surrogate_max_spread=0.05, seed=1000 + seed, **kw)
        flat = abi.sqa_chi_run(*args, chi_override=np.ones(cfg.scan_points), seed=1000 + seed, **kw)
        for r in (pil, sr, flat):
            assert r["best_energy"] == prob.energy(r["best_state"])[0]
        gap = lambda e: (e - pil["best_energy"]) / abs(pil["best_energy"])
        rows.append((fam, n, beta, bool(sr["surrogate_used"]), float(sr["surrogate_spread"]), gap(sr["best_energy"]),
                     gap(flat["best_energy"]), pil["kernel_ms"], sr["kernel_ms"]))
        print(f"{fam} n={n} beta={beta}: used={rows[-1][3]} spread={rows[-1][4]:.3f} gap surrogate "
              f"{rows[-1][5]:+.3%} flat {rows[-1][6]:+.3%} ms pilot-run {pil['kernel_ms']:.0f} "
              f"surrogate-run {sr['kernel_ms']:.0f}", flush=True)
    used = [r[3] for r in rows]
    g_sur = np.array([r[5] for r in rows])
    g_flat = np.array([r[6] for r in rows])
    ms_p, ms_s = sum(r[7] for r in rows), sum(r[8] for r in rows)
    print(f"mean gap surrogate {g_sur.mean():+.3%} flat {g_flat.mean():+.3%}; device time pilot-scheduled "
          f"{ms_p:.0f} ms, surrogate-scheduled {ms_s:.0f} ms ({1 - ms_s / ms_p:.1%} saved)", flush=True)
    assert all(used), "the surrogate fell back to the pilot on a held-out instance"
    assert g_sur.mean() <= 0.01
    assert ms_s <= 0.90 * ms_p
