#!/usr/bin/env python3
"""Throughput of the production SQA sweep on one MI355X.

Workload (one timed step = one full anneal): plain Trotterized SQA on a dense SK Ising, N=512,
J_ij ~ N(0, 1/N) symmetric with zero diagonal, P=32 Trotter slices, 256 replicas, beta=32, Gamma
linear 3.0 -> 1e-3 over 2000 sweeps (one sweep per schedule step), FP32 local fields, seeded
synthetic instance. Attempts per step = 2000 * 256 * 32 * 512 = 8.39e9 single-spin Metropolis
attempts.

Prints one JSON line: whole-job attempts/s, the kernel's HBM-roofline fraction (algorithmic bytes
10.25 B/attempt: fields read+write 8 B, spins read+write 2 B, J once per sweep 0.25 B, over the
hipEvent-timed kernel duration), and the reference's CPU path timed on this host's cores over a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import re
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, P, R, SWEEPS = 512, 32, 256, 2000
BETA, G0, G1 = 32.0, 3.0, 1e-3
ATTEMPTS = SWEEPS * R * P * N
# algorithmic bytes per sweep: fields r+w (2*R*P*N*4) + spins r+w (2*R*P*N) + J once (N*N*4)
BYTES_PER_SWEEP = 2 * R * P * N * 4 + 2 * R * P * N + N * N * 4
BYTES_PER_LAUNCH = BYTES_PER_SWEEP * SWEEPS
HBM_PEAK_GBS = 8000.0
METRIC = "spin-flip attempts/s; time-to-99%-target energy vs plain GPU SQA, 1 GPU"


def instance(seed: int = 20261203):
    rng = np.random.default_rng(seed)
    J = rng.normal(0.0, 1.0 / np.sqrt(N), (N, N))
    J = np.triu(J, 1)
    return J + J.T


def cpu_baseline(J, betas, gammas, sample_steps: int):
    """Reference build (or the oracle restatement when it is absent) on the host cores."""
    from oracle.orc import Oracle, Problem, REF_PORTABLE_SO, ORACLE_SO

    path, kind = (REF_PORTABLE_SO, "reference") if os.path.exists(REF_PORTABLE_SO) else (ORACLE_SO, "port")
    cores = min(16, os.cpu_count() or 1)
    lib = Oracle(path)
    lib.set_threads(cores)
    p = Problem.dense(np.zeros(N), J)
    t = time.perf_counter()
    lib.sqa_run(p, betas[:sample_steps], gammas[:sample_steps], P, R, 1, 0, 0, 0, 1)
    dt = time.perf_counter() - t
    att = sample_steps * R * P * N
    return {
        "value": att / dt, "unit": "attempts/s", "cores": cores, "kind": kind,
        "sample": (f"same N=512 P=32 R=256 instance, first {sample_steps} of {SWEEPS} schedule steps "
                   f"(slice sweeps + the reference's serial per-step FP64 slice energies), {dt:.1f} s"),
    }


def traffic_from_profile():
    """HBM bytes per sweep launch from the committed counter summary (FETCH_SIZE + WRITE_SIZE
    passes, each calibrated on a copy kernel of known size), or None."""
    f = os.path.join(ROOT, "profiles", "sqa_pipe_kernel_c2_traffic.md")
    if not os.path.exists(f):
        return None
    m = re.search(r"^hbm_bytes_per_launch:\s*(\d+)", open(f).read(), re.M)
    return int(m.group(1)) if m else None


def _line(metric_value, steps, warmup, wall, workload_cfg, roofline, dtype="f32",
          data="synthetic (seeded SK instance, J_ij ~ N(0, 1/N))"):
    return {
        "metric": METRIC, "value": metric_value, "unit": "attempts/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": data, "config": workload_cfg, "roofline": roofline,
    }


def sk_instance(n: int, seed: int):
    rng = np.random.default_rng(seed)
    J = rng.normal(0.0, 1.0 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def bench_c1(args):
    """Plain SA on a dense SK Ising: N=64, 32 reads, geometric beta 0.1 -> 3.0 over 100 steps x 10
    sweeps; all 32 reads in one launch (Annealer.run_reads). Not HBM-bound (a 16 KB FP32 J): the
    roofline entry states the algorithmic bytes (J row 256 B read per sweep-site step of the 32-read
    group, fields r+w 8 B and spins r+w 2 B per attempt) against HBM peak for completeness."""
    import qlatanneal_amd as qa
    n, reads, steps, spb = 64, 32, 100, 10
    J = sk_instance(n, 20261203)
    betas = (0.1 * (30.0 ** (np.arange(steps) / (steps - 1)))).tolist()
    ann = qa.Annealer(qa.DenseIsing(np.zeros(n), J), qa.AnnealSchedule.from_betas(betas), backend="hip")
    attempts = reads * steps * spb * n
    for w in range(args.warmup):
        ann.run_reads(spb, reads, 10 + w)
    t0 = time.perf_counter()
    best = []
    for k in range(args.steps):
        best.append(min(r.best_energy for r in ann.run_reads(spb, reads, 1000 + 32 * k)))
    wall = time.perf_counter() - t0
    bytes_per_launch = steps * spb * n * (n * 4 + reads * 10)
    sec = wall / args.steps
    achieved = bytes_per_launch / sec / 1e9
    line = _line(args.steps * attempts / wall, args.steps, args.warmup, wall,
                 {"workload": "sa_dense_sk_n64_32reads_100x10_geometric_fp32", "N": n, "reads": reads,
                  "steps": steps, "sweeps_per_beta": spb, "attempts_per_step": attempts,
                  "best_energy_mean": float(np.mean(best)), "timing": "wall clock of run_reads (launch + copies)"},
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None})
    if not args.no_cpu:
        from oracle.orc import Oracle, Problem, REF_PORTABLE_SO, ORACLE_SO
        path, kind = (REF_PORTABLE_SO, "reference") if os.path.exists(REF_PORTABLE_SO) else (ORACLE_SO, "port")
        lib = Oracle(path)
        lib.set_threads(1)
        p = Problem.dense(np.zeros(n), J)
        t = time.perf_counter()
        for r in range(reads):  # solve() runs reads one after another (solver.py:679)
            lib.sa_run(p, np.array(betas), spb, 1000 + r)
        dt = time.perf_counter() - t
        line["cpu_baseline"] = {"value": attempts / dt, "unit": "attempts/s", "cores": 1, "kind": kind,
                                "sample": f"the whole workload (32 sequential reads), {dt:.2f} s"}
    print(json.dumps(line), flush=True)


def bench_c3(args):
    """SQA-PT with matrix-core field refresh: dense SK N=2048, P=64, 32 rungs x 4 ladders, 10 sweeps
    per step, 500 steps (5000 sweeps), swaps every step, FP32 fields refreshed every 100 sweeps (bf16
    hi+lo S.J, FP32 accumulate) with the FP64 drift check. Algorithmic bytes per sweep (128
    replicas): fields r+w 134.2 MB + spins r+w 33.6 MB + J once 16.8 MB = 11.0 B/attempt; the
    timed device span includes the swap and refresh kernels."""
    from qlatanneal_amd import abi
    n, P, rungs, L, spb, steps = 2048, 64, 32, 4, 10, 500
    J = sk_instance(n, 20261204)
    betas = np.geomspace(0.3, 5.0, rungs)
    gammas = np.geomspace(3.0, 0.05, rungs)
    ctx = abi.Context(0)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    attempts = L * rungs * P * n * spb * steps
    for w in range(args.warmup):
        abi.sqapt_run(dp, betas, gammas, P, spb, 0, 5, 1, 0, L, abi.MODE_FP32, 10 + w, refresh_interval=100)
    kms, drift, best = [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = abi.sqapt_run(dp, betas, gammas, P, spb, 0, steps, 1, 0, L, abi.MODE_FP32, 1000 + k,
                          refresh_interval=100)
        kms.append(r["kernel_ms"])
        drift.append(r["field_drift"].tolist())
        best.append(float(np.min(r["best_energy"])))
        print(f"step {k}: device {kms[-1]:.1f} ms best {best[-1]:.4f} drift {drift[-1]}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    sweeps = spb * steps
    bytes_per_launch = (2 * L * rungs * P * n * 4 + 2 * L * rungs * P * n + n * n * 4) * spb
    launch_s = float(np.mean(kms)) / 1e3 / steps
    achieved = bytes_per_launch / launch_s / 1e9
    line = _line(args.steps * attempts / wall, args.steps, args.warmup, wall,
                 {"workload": "sqapt_dense_sk_n2048_p64_32rungs_x4_5000sweeps_refresh100_fp32", "N": n,
                  "trotter_slices": P, "rungs": rungs, "ladders": L, "sweeps": sweeps,
                  "attempts_per_step": attempts, "device_ms_mean": float(np.mean(kms)),
                  "field_drift_before_after_refresh": drift[-1], "best_energy_per_spin": float(np.min(best)) / n},
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None})
    if not args.no_cpu:
        from oracle.orc import Oracle, Problem, REF_PORTABLE_SO, ORACLE_SO
        path, kind = (REF_PORTABLE_SO, "reference") if os.path.exists(REF_PORTABLE_SO) else (ORACLE_SO, "port")
        cores = min(16, os.cpu_count() or 1)
        lib = Oracle(path)
        lib.set_threads(cores)
        p = Problem.dense(np.zeros(n), J)
        ss = max(1, args.cpu_sample_steps // 8)
        t = time.perf_counter()
        lib.sqapt_run(p, betas, gammas, P, spb, 0, ss, 1, 0, 0, 7)  # one ladder
        dt = time.perf_counter() - t
        att = rungs * P * n * spb * ss
        line["cpu_baseline"] = {"value": att / dt, "unit": "attempts/s", "cores": cores, "kind": kind,
                                "sample": (f"one ladder (32 rungs), {ss} of {steps} steps of the same instance "
                                           f"({dt:.1f} s); the reference's cost per step is acceptance-independent")}
    print(json.dumps(line), flush=True)


def maxcut3_edges(n: int, seed: int):
    """3-regular-style MAX-CUT graph with unit weights: edges from random perfect matchings."""
    rng = np.random.default_rng(seed)
    edges = set()
    while len(edges) < 3 * n // 2:
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            u, v = sorted((int(perm[i]), int(perm[i + 1])))
            if len(edges) < 3 * n // 2:
                edges.add((u, v))
    return [(u, v, 1.0) for u, v in sorted(edges)]


def bench_c5(args):
    """sqa_chi at the C5 shape without a surrogate (the pilot-scheduled method the surrogate
    replaces): run_chi on a dense SK N=2048 and a 3-regular MAX-CUT N=4096, P=32, R=4, beta=4,
    Gamma 3.0 -> 1e-3, pilot of 64 points x (10 burn + 30 measured) sweeps, production 100 steps x
    10 sweeps. One timed step = both runs. Algorithmic bytes: ~10 B per attempt (a chain's field and
    spin read and written), as for the other SQA lines; the sweeps are latency-bound."""
    from qlatanneal_amd import abi
    P, R, beta, pts, burn, meas, steps, spb = 32, 4, 4.0, 64, 10, 30, 100, 10
    ctx = abi.Context(0)
    n_sk, n_mc = 2048, 4096
    J = sk_instance(n_sk, 20261208)
    probs = [("sk", n_sk, abi.DeviceProblem.dense(ctx, np.zeros(n_sk), J)),
             ("maxcut3", n_mc, abi.DeviceProblem.sparse(ctx, np.zeros(n_mc), maxcut3_edges(n_mc, 20261209), 0.0))]
    kw = dict(num_steps=steps, sweeps_per_step=spb, scan_points=pts, scan_sweeps=meas, scan_burn=burn)
    attempts = sum(R * P * n * (pts * (burn + meas) + steps * spb) for _, n, _ in probs)
    for w in range(args.warmup):
        for _, _, dp in probs:
            abi.sqa_chi_run(dp, P, R, beta, 3.0, 1e-3, num_steps=4, sweeps_per_step=2, scan_points=4, scan_sweeps=2,
                            scan_burn=1, seed=10 + w)
    kms, best = [], {}
    t0 = time.perf_counter()
    for k in range(args.steps):
        tot = 0.0
        for name, n, dp in probs:
            r = abi.sqa_chi_run(dp, P, R, beta, 3.0, 1e-3, seed=1000 + 97 * k, **kw)
            tot += r["kernel_ms"]
            best.setdefault(name, []).append(r["best_energy"] / n)
        kms.append(tot)
        print(f"step {k}: run_chi device {tot:.1f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    kern_s = float(np.mean(kms)) / 1e3
    achieved = attempts * 10 / kern_s / 1e9
    line = _line(args.steps * attempts / wall, args.steps, args.warmup, wall,
                 {"workload": "sqa_chi_run_sk_n2048_and_maxcut3_n4096_p32_r4_pilot64x40_prod100x10_fp32",
                  "trotter_slices": P, "replicas": R, "beta": beta, "scan_points": pts,
                  "pilot_sweeps": burn + meas, "production": [steps, spb], "attempts_per_step": attempts,
                  "device_ms_mean": float(np.mean(kms)),
                  "best_energy_per_spin": {k: float(np.min(v)) for k, v in best.items()}},
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None},
                 data="synthetic (seeded SK J_ij ~ N(0, 1/N); 3-regular unit-weight MAX-CUT)")
    if not args.no_cpu:
        from oracle.orc import Oracle, Problem, REF_PORTABLE_SO, ORACLE_SO
        path, kind = (REF_PORTABLE_SO, "reference") if os.path.exists(REF_PORTABLE_SO) else (ORACLE_SO, "port")
        cores = min(16, os.cpu_count() or 1)
        lib = Oracle(path)
        lib.set_threads(cores)
        p = Problem.dense(np.zeros(n_sk), J)
        sp, sb, sm, st, ss = 4, 1, 2, 20, 2
        t = time.perf_counter()
        lib.sqa_chi_run(p, P, R, beta, 3.0, 1e-3, num_steps=st, sweeps_per_step=ss, scan_points=sp, scan_sweeps=sm,
                        scan_burn=sb, seed=7)
        dt = time.perf_counter() - t
        att = R * P * n_sk * (sp * (sb + sm) + st * ss)
        line["cpu_baseline"] = {"value": att / dt, "unit": "attempts/s", "cores": cores, "kind": kind,
                                "sample": (f"run_chi on the same SK N=2048 instance with a {sp}-point pilot of "
                                           f"{sb}+{sm} sweeps and {st} x {ss} production sweeps ({dt:.1f} s)")}
    print(json.dumps(line), flush=True)


class _Hubo:
    """A HigherOrderIsing plus its (already consolidated: distinct sorted triplets) terms."""

    def __init__(self, ham, trips, coeffs):
        self.ham, self.trips, self.coeffs = ham, trips, coeffs


def c4_sample(count: int, seed: int = 20261207):
    """A seeded sample of the C4 instance mix: 80 % dense QUBOs (N ~ U[50, 1500], edge density
    ~ U[0.05, 1], integer Q_ij ~ U{-100..100} symmetric on the kept pairs, converted with
    QUBO.to_ising) and 20 % 3-body HUBOs (N = 256, 4N distinct triplets, J ~ N(0, 1))."""
    import qlatanneal_amd as qa
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        if k % 5 == 4:
            n = 256
            trip = set()
            while len(trip) < 4 * n:
                trip.add(tuple(sorted(rng.choice(n, 3, replace=False).tolist())))
            trips = sorted(trip)
            coeffs = [float(rng.normal()) for _ in trips]
            hobj = qa.HigherOrderIsing([(list(t), c) for t, c in zip(trips, coeffs)], n)
            out.append(("hubo", n, _Hubo(hobj, trips, coeffs)))
        else:
            n = int(rng.integers(50, 1501))
            dens = rng.uniform(0.05, 1.0)
            Q = rng.integers(-100, 101, (n, n)).astype(float)
            keep = rng.random((n, n)) < dens
            Q = np.triu(Q * keep)
            Q = Q + np.triu(Q, 1).T
            out.append(("qubo", n, qa.QUBO(Q).to_ising()))
    return out


def bench_c4(args):
    """C4's instances through the plain per-instance GPU path (the batched persistent launch is
    switched off): SQA P=16, 64 replicas per instance (the 64 reads of an instance run together),
    auto_schedule_sqa (50 steps x 20 slice sweeps + 5 worldline sweeps). One timed step = a
    seeded sample of `--c4-instances` instances of the C4 mix, driven through the SQAAnnealer
    class by `--c4-streams` host threads, largest instance first (1: one after another);
    algorithmic bytes ~10 B per attempt (fields and spins read and written) plus J once per
    instance sweep."""
    import qlatanneal_amd as qa
    P, R, spb, wl = 16, 64, 20, 5
    sched = qa.auto_schedule_sqa()
    steps = len(sched.betas)
    inst = c4_sample(args.c4_instances)
    attempts = sum(steps * R * (spb * P * n + wl * n) for _, n, _ in inst)
    nbytes = sum(steps * (spb + wl) * (R * P * n * 10 + n * n * 4) for _, n, _ in inst)
    kms, best = [], []
    # K host threads take the instances largest first; each run leases a free one of the library's
    # HIP contexts (one stream each), so with K > 1 runs execute concurrently on the GPU's hardware
    # queues (the bindings release the GIL during a run).
    anns = [qa.SQAAnnealer(h.ham if kind == "hubo" else h, sched, P, R, backend="hip") for kind, _, h in inst]
    K = max(1, args.c4_streams)
    order = sorted(range(len(inst)), key=lambda i: -inst[i][1])

    def one(i, seed0):
        anns[i].set_seed(seed0 + i)
        return anns[i].run(spb, wl).best_energy, anns[i].last_kernel_ms()

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(K) as pool:
        for w in range(args.warmup):
            list(pool.map(lambda i: one(i, 10 + w), order[:K]))
        t0 = time.perf_counter()
        for k in range(args.steps):
            res = list(pool.map(lambda i: one(i, 1000 + 97 * k), order))
            best.extend(e for e, _ in res)
            kms.append(sum(m for _, m in res))
            print(f"step {k}: {K} thread(s), sweep kernels {kms[-1]:.1f} ms summed over {len(anns)} instances",
                  file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
    kern_s = float(np.mean(kms)) / 1e3 if K == 1 else wall / args.steps
    achieved = nbytes / kern_s / 1e9
    sizes = sorted(n for _, n, _ in inst)
    line = _line(args.steps * attempts / wall, args.steps, args.warmup, wall,
                 {"workload": f"c4_mix_sample{len(inst)}_sqa_p16_r64_auto50x(20+5)_fp32_per_instance",
                  "instances": len(inst), "hubo": sum(1 for k, _, _ in inst if k == "hubo"),
                  "n_min": sizes[0], "n_median": sizes[len(sizes) // 2], "n_max": sizes[-1],
                  "attempts_per_step": attempts, "sweep_kernel_ms_mean": kern_s * 1e3,
                  "streams": K,
                  "path": ("SQAAnnealer per instance (the batched launch is switched off)"
                           + ("" if K == 1 else f", {K} host threads over the library's HIP streams"))},
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None},
                 data="synthetic (seeded C4 mix: integer QUBOs of random density and 3-body HUBOs)")
    if not args.no_cpu:
        from oracle.orc import Oracle, Problem, REF_PORTABLE_SO, ORACLE_SO
        path, kind = (REF_PORTABLE_SO, "reference") if os.path.exists(REF_PORTABLE_SO) else (ORACLE_SO, "port")
        lib = Oracle(path)
        lib.set_threads(1)  # one read of one replica: the reference's OpenMP fork is over replicas
        _, n, h = next(x for x in inst if x[0] == "qubo")
        p = Problem.dense(np.asarray(h.fields()), np.asarray(h.couplings()), h.constant())
        nreads = 4
        t = time.perf_counter()
        for r in range(nreads):  # the reference runs an instance's reads one after another
            lib.sqa_run(p, np.array(sched.betas), np.array(sched.gammas), P, 1, spb, wl, 0, 0, 3 + r)
        dt = time.perf_counter() - t
        att = nreads * steps * (spb * P * n + wl * n)
        line["cpu_baseline"] = {"value": att / dt, "unit": "attempts/s", "cores": 1, "kind": kind,
                                "sample": (f"{nreads} of the 64 reads (1 replica each, the full schedule) of the "
                                           f"first QUBO instance (N={n}), {dt:.1f} s")}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("c2", "c1", "c3", "c4", "c5"), default="c2",
                    help="c2 (default): the headline SQA sweep; c1: SA reads; c3: SQA-PT with field refresh; "
                         "c4: a sample of the batched-QUBO mix through the per-instance path")
    ap.add_argument("--c4-instances", type=int, default=20)
    ap.add_argument("--c4-streams", type=int, default=4,
                    help="c4: host threads driving the instances (each annealer runs on one of the "
                         "library's HIP streams; 1 runs the instances one after another)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every run here has a single MI355X, and the sweep shards by "
              "independent replicas and instances, one GPU each; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c1":
        bench_c1(args)
        return
    if args.workload == "c3":
        bench_c3(args)
        return
    if args.workload == "c5":
        bench_c5(args)
        return
    if args.workload == "c4":
        bench_c4(args)
        return

    import qlatanneal_amd as qa

    J = instance()
    betas = np.full(SWEEPS, BETA)
    gammas = np.linspace(G0, G1, SWEEPS)
    ham = qa.DenseIsing(np.zeros(N), J, 0.0)
    sched = qa.SQASchedule.from_vectors(betas.tolist(), gammas.tolist())
    ann = qa.SQAAnnealer(ham, sched, P, R, backend="hip")
    for w in range(args.warmup):
        ann.set_seed(100 + w)
        ann.run(1, 0)
        print(f"warmup {w} done", file=sys.stderr, flush=True)
    kms = []
    best = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ann.set_seed(1000 + k)
        res = ann.run(1, 0)  # blocks until the device run has completed
        kms.append(ann.last_kernel_ms())
        best.append(res.best_energy)
        print(f"step {k}: kernel {kms[-1]:.1f} ms best {res.best_energy:.4f}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    acc = ann.last_accept_rate()
    value = args.steps * ATTEMPTS / wall
    kern_s = float(np.mean(kms)) / 1e3
    achieved = BYTES_PER_LAUNCH / kern_s / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SK instance, J_ij ~ N(0, 1/N))",
        "config": {"workload": "sqa_dense_sk_n512_p32_r256_2000sweeps_fp32", "N": N, "trotter_slices": P,
                   "replicas": R, "sweeps": SWEEPS, "beta": BETA, "gamma": [G0, G1],
                   "attempts_per_step": ATTEMPTS, "accept_rate": acc,
                   "kernel_ms_mean": kern_s * 1e3, "best_energy_per_spin_mean": float(np.mean(best)) / N},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile()},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(J, betas, gammas, args.cpu_sample_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
