"""Learned chi_B surrogate for SQAChiAnnealer.run_chi (no counterpart in the reference).

run_chi (sqa_chi_annealer.cpp:232-281) spends a pilot scan -- scan_points x (burn + sweeps)
checkerboard sweeps -- to measure the bond susceptibility chi_B on a uniform s grid, and builds the
production Gamma(t) schedule from the normalized profile chi_B / max chi_B alone (peak, relative
floor, cumulative weight; :282-340). This module learns that profile from instance statistics:

  features  qa_chi_features: size, couplings per spin, field scale and heterogeneity, coupling
            kurtosis and sign balance, |h| ratio, interaction order, beta, P, A0, Gamma range;
  labels    normalized pilot profiles measured by the library's own FP32 pilot kernel (qa_chi_pilot);
  network   an ensemble of MLPs (default 4 members of 12 -> 256 x 3 -> 64) trained and evaluated by
            the library's hand-written kernels (qa_mlp_*: bf16 weights, FP32 accumulation);
  use       qa_sqa_chi_run(surrogate=...) takes the ensemble mean as the profile, or runs the real
            pilot when the ensemble spread is above a bound or the instance lies outside the
            training range of the features.

Training families follow the sqa_chi workload: Sherrington-Kirkpatrick (J_ij ~ N(0, 1/N)) and
3-regular MAX-CUT with unit antiferromagnetic couplings (J = +1 on edges, which the energy
E = sum J s_i s_j minimizes by cutting edges).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import abi

NFEAT = abi.CHI_NFEATURES


@dataclass
class PilotConfig:
    """The run_chi arguments a surrogate is trained for (the pilot grid is part of its output)."""
    trotter_slices: int = 32
    replicas: int = 4
    gamma_start: float = 3.0
    gamma_end: float = 1e-3
    driver_A0: float = 3.0
    scan_points: int = 64
    scan_sweeps: int = 30
    scan_burn: int = 10


def sk_problem(ctx: abi.Context, rng: np.random.Generator, n: int) -> abi.DeviceProblem:
    J = rng.normal(0.0, 1.0 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return abi.DeviceProblem.dense(ctx, np.zeros(n), J + J.T, 0.0)


def regular3_edges(rng: np.random.Generator, n: int) -> List[Tuple[int, int]]:
    """A uniformly paired 3-regular simple graph on n (even) vertices (pairing model with restarts)."""
    if n % 2 or n < 4:
        raise ValueError("3-regular graphs need an even n >= 4")
    for _ in range(1000):
        pts = rng.permutation(np.repeat(np.arange(n), 3))
        a, b = pts[0::2], pts[1::2]
        if np.any(a == b):
            continue
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        if np.unique(lo * n + hi).size != lo.size:
            continue
        return list(zip(lo.tolist(), hi.tolist()))
    raise RuntimeError("could not draw a simple 3-regular graph")


def maxcut3_problem(ctx: abi.Context, rng: np.random.Generator, n: int) -> abi.DeviceProblem:
    edges = [(i, j, 1.0) for i, j in regular3_edges(rng, n)]
    return abi.DeviceProblem.sparse(ctx, np.zeros(n), edges, 0.0)


FAMILIES = {"sk": sk_problem, "maxcut3": maxcut3_problem}


@dataclass
class InstanceSpec:
    family: str
    n: int
    beta: float
    seed: int


def random_specs(count: int, rng: np.random.Generator, sk_n=(256, 2048), maxcut_n=(256, 4096),
                 beta=(1.0, 16.0)) -> List[InstanceSpec]:
    """Training/validation instances: half SK, half 3-regular MAX-CUT, log-uniform N and beta."""
    out = []
    for k in range(count):
        fam = "sk" if k % 2 == 0 else "maxcut3"
        lo, hi = sk_n if fam == "sk" else maxcut_n
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        n += n % 2
        b = float(np.exp(rng.uniform(np.log(beta[0]), np.log(beta[1]))))
        out.append(InstanceSpec(fam, n, b, int(rng.integers(1, 2**62))))
    return out


def build(ctx: abi.Context, spec: InstanceSpec) -> abi.DeviceProblem:
    return FAMILIES[spec.family](ctx, np.random.default_rng(spec.seed), spec.n)


def features(prob: abi.DeviceProblem, beta: float, cfg: PilotConfig) -> np.ndarray:
    return abi.chi_features(prob, beta, cfg.trotter_slices, cfg.driver_A0, cfg.gamma_start, cfg.gamma_end)


def pilot_profile(prob: abi.DeviceProblem, beta: float, cfg: PilotConfig, seed: int):
    """(normalized profile chi / max chi, raw pilot dict); the profile is None for a flat scan."""
    r = abi.chi_pilot(prob, cfg.trotter_slices, cfg.replicas, beta, cfg.gamma_start, cfg.gamma_end,
                      cfg.scan_points, cfg.scan_sweeps, cfg.scan_burn, cfg.driver_A0, seed)
    mx = float(np.max(r["scan_chi_B"]))
    prof = r["scan_chi_B"] / mx if mx > 0 and np.isfinite(mx) else None
    return prof, r


def make_dataset(ctx: abi.Context, specs: Sequence[InstanceSpec], cfg: PilotConfig, log=None, threads: int = 1):
    """Features X [S][NFEAT] and labels Y [S][scan_points] from GPU pilot scans (flat scans dropped),
    plus the pilot device time in ms. With threads > 1 the scans run on that many contexts (one HIP
    stream each) at once; the dataset is the same, in spec order."""
    def one(c: abi.Context, sp: InstanceSpec):
        prob = build(c, sp)
        prof, r = pilot_profile(prob, sp.beta, cfg, sp.seed ^ 0x5DEECE66D)
        x = features(prob, sp.beta, cfg) if prof is not None else None
        del prob
        return x, prof, r["kernel_ms"]

    X, Y, ms = [], [], 0.0
    step = max(1, len(specs) // 10)
    if threads <= 1:
        results = (one(ctx, sp) for sp in specs)
    else:
        import threading
        from concurrent.futures import ThreadPoolExecutor
        local = threading.local()

        def task(sp):
            if not hasattr(local, "ctx"):
                local.ctx = abi.Context(ctx.device)
            return one(local.ctx, sp)

        pool = ThreadPoolExecutor(threads)
        results = pool.map(task, specs)
    try:
        for k, (x, prof, t) in enumerate(results):
            ms += t
            if prof is not None:
                X.append(x)
                Y.append(prof)
            if log and (k + 1) % step == 0:
                log(f"pilot scans {k + 1}/{len(specs)}")
    finally:
        if threads > 1:
            pool.shutdown()
    return np.array(X).reshape(-1, NFEAT), np.array(Y).reshape(-1, cfg.scan_points), ms


class ChiSurrogate:
    """Ensemble MLP predicting normalized pilot chi_B profiles (device training and inference)."""

    def __init__(self, ctx: abi.Context, cfg: Optional[PilotConfig] = None, width: int = 256, depth: int = 3,
                 members: int = 4, seed: int = 0):
        self.ctx = ctx
        self.cfg = cfg or PilotConfig()
        self.mlp = abi.Mlp(ctx, NFEAT, width, depth, self.cfg.scan_points, members, seed)

    def fit(self, X, Y, epochs: int = 400, batch: int = 64, lr: float = 2e-3, weight_decay: float = 1e-5,
            seed: int = 1) -> np.ndarray:
        return self.mlp.train(X, Y, epochs, batch, lr, weight_decay, seed)

    def predict(self, X) -> Tuple[np.ndarray, np.ndarray]:
        """(ensemble mean [S][scan_points], spread [S]: largest per-point std across members)."""
        y = self.mlp.forward(X).astype(np.float64)
        return y.mean(axis=0), y.std(axis=0).max(axis=1)

    def save(self, path: str) -> None:
        params, norm = self.mlp.get_state()
        c = self.cfg
        np.savez(path, params=params, norm=norm,
                 dims=np.array([self.mlp.n_in, self.mlp.width, self.mlp.depth, self.mlp.n_out, self.mlp.members]),
                 cfg=np.array([c.trotter_slices, c.replicas, c.gamma_start, c.gamma_end, c.driver_A0, c.scan_points,
                               c.scan_sweeps, c.scan_burn], np.float64))

    @classmethod
    def load(cls, ctx: abi.Context, path: str) -> "ChiSurrogate":
        with np.load(path, allow_pickle=False) as z:
            n_in, width, depth, n_out, members = (int(v) for v in z["dims"])
            c = z["cfg"]
            cfg = PilotConfig(int(c[0]), int(c[1]), float(c[2]), float(c[3]), float(c[4]), int(c[5]), int(c[6]),
                              int(c[7]))
            if n_in != NFEAT or n_out != cfg.scan_points:
                raise ValueError("surrogate file does not match the feature/scan layout")
            s = cls(ctx, cfg, width, depth, members)
            s.mlp.set_state(z["params"], z["norm"])
        return s


def scan_grid(cfg: PilotConfig) -> Tuple[np.ndarray, np.ndarray]:
    """The pilot's s grid (uniform over [s(gamma_start), s(gamma_end)], s = A0 / (A0 + Gamma)) and its
    Gamma values, as run_chi lays them out (sqa_chi_annealer.cpp:232-281)."""
    A0 = cfg.driver_A0 if cfg.driver_A0 > 0 else cfg.gamma_start
    s_lo, s_hi = A0 / (A0 + cfg.gamma_start), A0 / (A0 + cfg.gamma_end)
    s = s_lo + (s_hi - s_lo) * np.arange(cfg.scan_points) / (cfg.scan_points - 1)
    return s, A0 * (1.0 - s) / s


def chi_profile(sur: Optional["ChiSurrogate"], prob: abi.DeviceProblem, beta: float, cfg: PilotConfig,
                max_spread: float = 0.05, seed: int = 0) -> Tuple[np.ndarray, bool, float]:
    """Normalized chi_B profile on the pilot grid: the surrogate's ensemble mean when the instance's
    features lie within a quarter width of the training range and the largest per-point member
    spread is at most max_spread, else the real pilot scan (qa_chi_pilot). Returns (profile,
    surrogate used, spread)."""
    if sur is not None:
        f = features(prob, beta, cfg)
        _, norm = sur.mlp.get_state()
        mean, scale, lo, hi = norm
        v = (f - mean) / scale
        w = 0.25 * (hi - lo)
        if np.all((v >= lo - w) & (v <= hi + w)):
            pred, spread = sur.predict(f[None, :])
            if spread[0] <= max_spread:
                return pred[0], True, float(spread[0])
    prof, r = pilot_profile(prob, beta, cfg, seed)
    return (prof if prof is not None else np.ones(cfg.scan_points)), False, float("nan")


def chi_ladder(profile: np.ndarray, cfg: PilotConfig, rungs: int, floor_fraction: float = 0.1) -> np.ndarray:
    """Gamma values of an SQA-PT ladder (hot to cold) at equal-weight quantiles of the chi_B profile
    over s: rungs crowd where chi_B is large (the quantum critical region, where neighbouring rungs
    must be close for exchanges to pass) and thin out where it is small. The weight is the
    normalized profile floored at floor_fraction, integrated by trapezoids on the pilot grid; the
    end rungs sit at gamma_start and gamma_end."""
    s, _ = scan_grid(cfg)
    w = np.maximum(np.asarray(profile, np.float64) / max(float(np.max(profile)), 1e-300), floor_fraction)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(s))])
    cdf /= cdf[-1]
    s_r = np.interp(np.linspace(0.0, 1.0, rungs), cdf, s)
    A0 = cfg.driver_A0 if cfg.driver_A0 > 0 else cfg.gamma_start
    g = A0 * (1.0 - s_r) / s_r
    return np.clip(g, cfg.gamma_end, cfg.gamma_start)


def chi_annealed_ladder(beta_schedule, gamma_schedule, rungs: int, beta_ratio: float = 0.5,
                        gamma_max: Optional[float] = None):
    """Per-step SQA-PT ladders that follow a run_chi schedule (qa_chi_schedule's beta and gamma per
    step): at step k every rung shares Gamma_k, and rung r runs at beta_k * beta_ratio**((R-1-r)/(R-1)),
    so rung R-1 anneals exactly on the schedule and the hotter rungs trail it by a geometric
    factor down to beta_ratio. With gamma_max the ladder starts at the first step whose Gamma is at
    most gamma_max (the steps before it are dropped). Returns (step_betas, step_gammas), each
    [steps, rungs], the qa_sqapt_params.step_betas / step_gammas layout."""
    if rungs < 2:
        raise ValueError("an SQA-PT ladder needs >= 2 rungs")
    if not 0.0 < beta_ratio <= 1.0:
        raise ValueError("beta_ratio must lie in (0, 1]")
    b = np.asarray(beta_schedule, np.float64)
    g = np.asarray(gamma_schedule, np.float64)
    if b.ndim != 1 or b.shape != g.shape:
        raise ValueError("beta_schedule and gamma_schedule must be 1-D of equal length")
    if gamma_max is not None:
        below = np.nonzero(g <= gamma_max)[0]
        if below.size == 0:
            raise ValueError("no step of the schedule has Gamma <= gamma_max")
        b, g = b[below[0]:], g[below[0]:]
    fr = beta_ratio ** ((rungs - 1 - np.arange(rungs)) / (rungs - 1))
    return b[:, None] * fr[None, :], np.repeat(g[:, None], rungs, axis=1)


def evaluate(ctx: abi.Context, sur: ChiSurrogate, specs: Sequence[InstanceSpec], num_steps: int = 100,
             sweeps_per_step: int = 10, max_spread: float = 0.05, seed: int = 12345, log=None) -> dict:
    """Production run_chi on each instance under four schedules, same production seed unless noted:
      pilot      the reference method (pilot scan, then production);
      surrogate  the surrogate's profile (pilot only on fallback);
      rerun      pilot-scheduled again with another seed (pilot and production): the run-to-run
                 spread a gap has to be read against;
      flat       no pilot and no surrogate: a constant profile, i.e. a schedule linear in s.
    Reports best-energy gaps relative to the pilot run, the fallback rate, profile errors
    (surrogate vs pilot, rerun pilot vs pilot) and device times."""
    cfg = sur.cfg
    rows = []
    for k, sp in enumerate(specs):
        prob = build(ctx, sp)
        kw = dict(num_steps=num_steps, sweeps_per_step=sweeps_per_step, scan_points=cfg.scan_points,
                  scan_sweeps=cfg.scan_sweeps, scan_burn=cfg.scan_burn, driver_A0=cfg.driver_A0)
        args = (prob, cfg.trotter_slices, cfg.replicas, sp.beta, cfg.gamma_start, cfg.gamma_end)
        pil = abi.sqa_chi_run(*args, seed=seed + k, **kw)
        sur_r = abi.sqa_chi_run(*args, surrogate=sur.mlp, surrogate_max_spread=max_spread, seed=seed + k, **kw)
        rer = abi.sqa_chi_run(*args, seed=seed + k + 7919, **kw)
        flat = abi.sqa_chi_run(*args, chi_override=np.ones(cfg.scan_points), seed=seed + k, **kw)
        f = features(prob, sp.beta, cfg)
        pred, spread = sur.predict(f[None, :])
        norm = lambda c: c / max(np.max(c), 1e-300)
        prof = norm(pil["scan_chi_B"])
        e_p = pil["best_energy"]
        gap = lambda e: (e - e_p) / max(abs(e_p), 1e-300)
        rows.append(dict(family=sp.family, n=sp.n, beta=sp.beta, e_pilot=e_p, e_surrogate=sur_r["best_energy"],
                         e_rerun=rer["best_energy"], e_flat=flat["best_energy"], gap=gap(sur_r["best_energy"]),
                         gap_rerun=gap(rer["best_energy"]), gap_flat=gap(flat["best_energy"]),
                         used=bool(sur_r["surrogate_used"]), spread=float(spread[0]),
                         ms_pilot_run=pil["kernel_ms"], ms_surrogate_run=sur_r["kernel_ms"],
                         profile_mae=float(np.mean(np.abs(pred[0] - prof))),
                         profile_mae_rerun=float(np.mean(np.abs(norm(rer["scan_chi_B"]) - prof))),
                         peak_s_err=float(abs(pil["scan_s"][np.argmax(pred[0])] - pil["s_star"]))))
        if log:
            r = rows[-1]
            log(f"eval {k + 1}/{len(specs)} {sp.family} n={sp.n} beta={sp.beta:.2f} gap={r['gap']:+.3%} "
                f"rerun={r['gap_rerun']:+.3%} flat={r['gap_flat']:+.3%} used={r['used']} "
                f"spread={r['spread']:.3f} mae={r['profile_mae']:.3f}/{r['profile_mae_rerun']:.3f} "
                f"ms pilot-run={r['ms_pilot_run']:.0f} surrogate-run={r['ms_surrogate_run']:.0f}")
        del prob
    col = lambda key: np.array([r[key] for r in rows], dtype=float)
    used = np.array([r["used"] for r in rows], dtype=bool)
    ms_p, ms_s = col("ms_pilot_run").sum(), col("ms_surrogate_run").sum()
    stats = lambda g: dict(mean=float(g.mean()), median=float(np.median(g)), max=float(g.max()),
                           within_1pct=float(np.mean(g <= 0.01)))
    return dict(instances=len(rows), surrogate_used_fraction=float(used.mean()),
                gap_surrogate=stats(col("gap")), gap_rerun=stats(col("gap_rerun")), gap_flat=stats(col("gap_flat")),
                profile_mae_surrogate=float(col("profile_mae").mean()),
                profile_mae_rerun=float(col("profile_mae_rerun").mean()),
                ms_pilot_scheduled=float(ms_p), ms_surrogate_scheduled=float(ms_s),
                wall_saving=float(1.0 - ms_s / ms_p) if ms_p > 0 else 0.0, rows=rows)


def _main(argv=None) -> int:
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description="Train or evaluate the chi_B surrogate on GPU pilot scans.")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--count", type=int, default=1200, help="pilot scans (instances) to generate")
    t.add_argument("--out", required=True, help="output .npz")
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--batch", type=int, default=64)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--members", type=int, default=4)
    t.add_argument("--seed", type=int, default=2024)
    t.add_argument("--data", default=None, help="also save the (X, Y) dataset here (.npz)")
    t.add_argument("--from-data", nargs="+", default=None,
                   help="train on these saved datasets (from `generate` or --data) instead of new scans")
    t.add_argument("--threads", type=int, default=1, help="concurrent pilot scans (one context each)")
    g = sub.add_parser("generate", help="pilot scans only, saved as a dataset chunk")
    g.add_argument("--count", type=int, required=True)
    g.add_argument("--seed", type=int, default=2024)
    g.add_argument("--chunk", type=int, default=0, help="chunk index: chunks of one seed draw disjoint instances")
    g.add_argument("--threads", type=int, default=4)
    g.add_argument("--data", required=True, help="output .npz (X, Y)")
    e = sub.add_parser("evaluate")
    e.add_argument("--model", required=True)
    e.add_argument("--count", type=int, default=16)
    e.add_argument("--seed", type=int, default=99)
    e.add_argument("--max-n", type=int, default=1024)
    e.add_argument("--num-steps", type=int, default=100)
    e.add_argument("--sweeps-per-step", type=int, default=10)
    e.add_argument("--max-spread", type=float, default=0.05)
    e.add_argument("--out", default=None, help="write the summary JSON here")
    a = ap.parse_args(argv)
    ctx = abi.Context(0)
    t0 = time.time()

    def log(msg):
        print(f"[{time.time() - t0:7.1f}s] {msg}", flush=True)

    if a.cmd == "generate":
        cfg = PilotConfig()
        specs = random_specs(a.count, np.random.default_rng([a.seed, a.chunk]))
        X, Y, ms = make_dataset(ctx, specs, cfg, log=log, threads=a.threads)
        np.savez(a.data, X=X, Y=Y)
        log(f"chunk {a.chunk}: {X.shape[0]} scans, pilot device time {ms / 1e3:.1f} s")
        print(json.dumps(dict(scans=int(X.shape[0]), pilot_ms=ms, wall_s=time.time() - t0)))
    elif a.cmd == "train":
        cfg = PilotConfig()
        if a.from_data:
            parts = []
            for path in a.from_data:
                with np.load(path, allow_pickle=False) as z:
                    parts.append((z["X"], z["Y"]))
            X = np.concatenate([x for x, _ in parts])
            Y = np.concatenate([y for _, y in parts])
            ms = float("nan")
        else:
            specs = random_specs(a.count, np.random.default_rng(a.seed))
            X, Y, ms = make_dataset(ctx, specs, cfg, log=log, threads=a.threads)
        log(f"dataset {X.shape[0]} scans, pilot device time {ms / 1e3:.1f} s")
        if a.data:
            np.savez(a.data, X=X, Y=Y)
        sur = ChiSurrogate(ctx, cfg, members=a.members, seed=a.seed)
        hist = sur.fit(X, Y, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed + 1)
        log(f"training loss first {hist[0]:.5f} last {hist[-1]:.5f}")
        sur.save(a.out)
        print(json.dumps(dict(scans=int(X.shape[0]), pilot_ms=ms, loss_first=float(hist[0]),
                              loss_last=float(hist[-1]), epochs=a.epochs)))
    else:
        sur = ChiSurrogate.load(ctx, a.model)
        rng = np.random.default_rng(a.seed)
        specs = random_specs(a.count, rng, sk_n=(256, a.max_n), maxcut_n=(256, a.max_n))
        res = evaluate(ctx, sur, specs, a.num_steps, a.sweeps_per_step, a.max_spread, log=log)
        summary = {k: v for k, v in res.items() if k != "rows"}
        print(json.dumps(summary))
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
