#!/usr/bin/env python3
"""Throughput of the annealing sweep path on one MI355X, and time-to-target against plain GPU SQA.

Default run (no --workload): the headline line is the largest single-GPU configuration that runs
here -- SQA-PT with matrix-core field refresh on a dense SK Ising, N=2048, P=64, 32 rungs x 4
ladders, 10 sweeps per PT step, an anneal of 500 PT steps, FP32 local fields refreshed every 100
sweeps with the FP64 drift check. A bench step is 25 PT steps (4.19e9 single-spin Metropolis
attempts); --steps 20 (the default) times exactly the configured 5000-sweep anneal. The same JSON
line carries
  roofline        the sweep kernel's algorithmic bytes per launch over its average launch time
                  against the 8 TB/s HBM peak, the counter-measured HBM bytes per launch, and the
                  lockstep critical path (lockstep steps per launch and measured time per step);
  cpu_baseline    the reference build (its sources compiled by oracle/Makefile) on all host cores
                  this process may use (nproc), on a bounded sample of the same workload;
  tts99           time-to-99%-target on N=2048 SK instances: FP32 sqa_chi on the chi_B surrogate's
                  schedule (no pilot scan) and batched FP32 SQA-PT whose ladders anneal on that
                  schedule (32 ladders per launch), runs of all instances concurrent, against the
                  plain per-instance GPU SQA path, every energy verified in FP64;
  other_workloads short timed runs of the other configurations (SA, plain SQA, the batched-QUBO
                  mix through the per-instance path, sqa_chi).
--workload c1 | c2 | c3 | c4 | c5 | tts runs one of them alone, with its own CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import subprocess
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

# Hardware queues of this process, set before anything initializes HIP. The concurrent parts of the
# bench (the C4 sample's runs on the library's 4 contexts; the tts99 methods' 4 host threads, each
# with a context of its own, beside the library's 4) open about ten streams; with HIP's default of
# 4 queues, two busy streams can land on one queue and run one after the other (measured: 16
# sqa_chi runs on 4 threads 7.2 s with 4 queues, 4.9 s with 16; one run alone 1.19 s). A larger
# value set by the caller is kept.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "spin-flip attempts/s; time-to-99%-target energy vs plain GPU SQA, 1 GPU"
SURROGATE_MODEL = os.path.join(ROOT, "qlatanneal_amd", "data", "chi_surrogate.npz")


def host_cores() -> int:
    """What `nproc` reports: the processing units this process may use."""
    try:
        return int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.strip())
    except Exception:
        return len(os.sched_getaffinity(0))


def cpu_lib(threads: int):
    """The reference build (its own sources, oracle/Makefile) when present, else the restatement."""
    from oracle.orc import Oracle, REF_PORTABLE_SO, ORACLE_SO
    path, kind = (REF_PORTABLE_SO, "reference") if os.path.exists(REF_PORTABLE_SO) else (ORACLE_SO, "port")
    lib = Oracle(path)
    lib.set_threads(threads)
    return lib, kind


def sk_instance(n: int, seed: int):
    rng = np.random.default_rng(seed)
    J = rng.normal(0.0, 1.0 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def traffic_from_profile(name: str):
    """HBM bytes per sweep launch from a committed counter summary (FETCH_SIZE and WRITE_SIZE
    passes, each scaled by the factor the same counter reads for a copy kernel of known size)."""
    f = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(f):
        return None
    m = re.search(r"^hbm_bytes_per_launch:\s*(\d+)", open(f).read(), re.M)
    return int(m.group(1)) if m else None


def _line(value, steps, warmup, wall, workload_cfg, roofline, dtype="f32",
          data="synthetic (seeded SK instance, J_ij ~ N(0, 1/N))"):
    return {
        "metric": METRIC, "value": value, "unit": "attempts/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": data, "config": workload_cfg, "roofline": roofline,
    }


def _roof(bytes_per_launch, launch_s, traffic=None, lockstep=None):
    achieved = bytes_per_launch / launch_s / 1e9
    r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}
    if lockstep is not None:
        # a workgroup advances its chains one lockstep step at a time: the launch cannot end before
        # its lockstep steps have run one after another
        r["critical_path"] = {"lockstep_steps_per_launch": lockstep, "ns_per_step": launch_s / lockstep * 1e9}
    return r


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---- C1: SA reads -----------------------------------------------------------------------------
def run_c1(steps, warmup, cpu=True):
    """Plain SA on a dense SK Ising: N=64, 32 reads, geometric beta 0.1 -> 3.0 over 100 steps x 10
    sweeps; all 32 reads in one launch (Annealer.run_reads). Not HBM-bound (a 16 KB FP32 J): the
    roofline entry states the algorithmic bytes (J row 256 B read per sweep-site step of the 32-read
    group, fields r+w 8 B and spins r+w 2 B per attempt) against HBM peak for completeness."""
    import qlatanneal_amd as qa
    n, reads, nb, spb = 64, 32, 100, 10
    J = sk_instance(n, 20261203)
    betas = (0.1 * (30.0 ** (np.arange(nb) / (nb - 1)))).tolist()
    ann = qa.Annealer(qa.DenseIsing(np.zeros(n), J), qa.AnnealSchedule.from_betas(betas), backend="hip")
    attempts = reads * nb * spb * n
    for w in range(warmup):
        ann.run_reads(spb, reads, 10 + w)
    t0 = time.perf_counter()
    best = []
    for k in range(steps):
        best.append(min(r.best_energy for r in ann.run_reads(spb, reads, 1000 + 32 * k)))
    wall = time.perf_counter() - t0
    bytes_per_launch = nb * spb * n * (n * 4 + reads * 10)
    line = _line(steps * attempts / wall, steps, warmup, wall,
                 {"workload": "sa_dense_sk_n64_32reads_100x10_geometric_fp32", "N": n, "reads": reads,
                  "steps": nb, "sweeps_per_beta": spb, "attempts_per_step": attempts,
                  "best_energy_mean": float(np.mean(best)), "timing": "wall clock of run_reads (launch + copies)"},
                 _roof(bytes_per_launch, wall / steps, lockstep=nb * spb * n))
    if cpu:
        from oracle.orc import Problem
        lib, kind = cpu_lib(1)
        p = Problem.dense(np.zeros(n), J)
        t = time.perf_counter()
        for r in range(reads):  # solve() runs reads one after another (solver.py:679)
            lib.sa_run(p, np.array(betas), spb, 1000 + r)
        dt = time.perf_counter() - t
        line["cpu_baseline"] = {"value": attempts / dt, "unit": "attempts/s", "cores": 1, "kind": kind,
                                "sample": f"the whole workload (32 sequential reads, as solve() runs them), {dt:.2f} s"}
    return line


# ---- C2: plain SQA ----------------------------------------------------------------------------
C2_N, C2_P, C2_R, C2_SWEEPS, C2_BETA = 512, 32, 256, 2000, 32.0


def run_c2(steps, warmup, cpu=True, cpu_sample_steps=8):
    """Plain Trotterized SQA on a dense SK Ising: N=512, P=32, 256 replicas, beta=32, Gamma linear
    3.0 -> 1e-3 over 2000 sweeps, FP32 local fields. Algorithmic bytes 10.25 B/attempt: fields r+w
    8 B, spins r+w 2 B, J once per sweep 0.25 B."""
    import qlatanneal_amd as qa
    N, P, R, SW = C2_N, C2_P, C2_R, C2_SWEEPS
    attempts = SW * R * P * N
    bytes_per_launch = (2 * R * P * N * 4 + 2 * R * P * N + N * N * 4) * SW
    J = sk_instance(N, 20261203)
    betas, gammas = np.full(SW, C2_BETA), np.linspace(3.0, 1e-3, SW)
    ham = qa.DenseIsing(np.zeros(N), J, 0.0)
    ann = qa.SQAAnnealer(ham, qa.SQASchedule.from_vectors(betas.tolist(), gammas.tolist()), P, R, backend="hip")
    for w in range(warmup):
        ann.set_seed(100 + w)
        ann.run(1, 0)
    kms, best = [], []
    t0 = time.perf_counter()
    for k in range(steps):
        ann.set_seed(1000 + k)
        res = ann.run(1, 0)  # blocks until the device run has completed
        kms.append(ann.last_kernel_ms())
        best.append(res.best_energy)
        log(f"c2 step {k}: kernel {kms[-1]:.1f} ms best {res.best_energy:.4f}")
    wall = time.perf_counter() - t0
    kern_s = float(np.mean(kms)) / 1e3
    line = _line(steps * attempts / wall, steps, warmup, wall,
                 {"workload": "sqa_dense_sk_n512_p32_r256_2000sweeps_fp32", "N": N, "trotter_slices": P,
                  "replicas": R, "sweeps": SW, "beta": C2_BETA, "gamma": [3.0, 1e-3],
                  "attempts_per_step": attempts, "accept_rate": ann.last_accept_rate(),
                  "kernel_ms_mean": kern_s * 1e3, "best_energy_per_spin_mean": float(np.mean(best)) / N},
                 _roof(bytes_per_launch, kern_s, traffic_from_profile("sqa_pipe_kernel_c2_traffic.md"),
                       lockstep=SW * (N + P - 1)))
    if cpu:
        line["cpu_baseline"] = c2_cpu_baseline(J, betas, gammas, cpu_sample_steps)
    return line


def c2_cpu_baseline(J, betas, gammas, sample_steps):
    """The CPU engine on the first `sample_steps` schedule steps of the C2 instance."""
    from oracle.orc import Problem
    N, P, R = C2_N, C2_P, C2_R
    cores = host_cores()
    lib, kind = cpu_lib(cores)
    p = Problem.dense(np.zeros(N), J)
    t = time.perf_counter()
    lib.sqa_run(p, betas[:sample_steps], gammas[:sample_steps], P, R, 1, 0, 0, 0, 1)
    dt = time.perf_counter() - t
    return {"value": sample_steps * R * P * N / dt, "unit": "attempts/s", "cores": cores, "kind": kind,
            "sample": (f"same N=512 P=32 R=256 instance, first {sample_steps} of {C2_SWEEPS} schedule steps "
                       f"(slice sweeps + the reference's serial per-step FP64 slice energies), {dt:.1f} s")}


def run_c2v(steps, warmup, cpu=True, sweeps=20, cpu_sample_steps=8):
    """C2's instance and shape (dense SK N=512, P=32, 256 replicas) in the FP64 validation mode
    (backend "hip_fp64": the reference's random draws and FP64 arithmetic, results bit-equal to
    the reference's), over the first `sweeps` steps of C2's schedule, one sweep each. Algorithmic
    bytes 18.5 B/attempt at N=512: FP64 fields r+w 16 B, spins r+w 2 B, FP64 J once per sweep 0.5 B."""
    import qlatanneal_amd as qa
    N, P, R = C2_N, C2_P, C2_R
    attempts = sweeps * R * P * N
    bytes_per_launch = (2 * R * P * N * 8 + 2 * R * P * N + N * N * 8) * sweeps
    J = sk_instance(N, 20261203)
    betas, gammas = np.full(C2_SWEEPS, C2_BETA), np.linspace(3.0, 1e-3, C2_SWEEPS)
    ham = qa.DenseIsing(np.zeros(N), J, 0.0)
    ann = qa.SQAAnnealer(ham, qa.SQASchedule.from_vectors(betas[:sweeps].tolist(), gammas[:sweeps].tolist()), P, R,
                         backend="hip_fp64")
    for w in range(warmup):
        ann.set_seed(100 + w)
        ann.run(1, 0)
    kms = []
    t0 = time.perf_counter()
    for k in range(steps):
        ann.set_seed(1000 + k)
        res = ann.run(1, 0)
        kms.append(ann.last_kernel_ms())
        log(f"c2v step {k}: kernel {kms[-1]:.1f} ms best {res.best_energy:.6f}")
    wall = time.perf_counter() - t0
    # outside the timed region: the returned best energy is the host FP64 energy of its state
    checked = ham.energy([int(x) for x in res.best_state.spins]) == res.best_energy
    kern_s = float(np.mean(kms)) / 1e3
    line = _line(steps * attempts / wall, steps, warmup, wall,
                 {"workload": f"sqa_dense_sk_n512_p32_r256_first{sweeps}sweeps_fp64_validation", "N": N,
                  "trotter_slices": P, "replicas": R, "sweeps": sweeps, "beta": C2_BETA,
                  "attempts_per_step": attempts, "kernel_ms_mean": kern_s * 1e3,
                  "best_energy_equals_fp64_recompute": bool(checked)},
                 _roof(bytes_per_launch, kern_s), dtype="f64")
    if cpu:
        line["cpu_baseline"] = c2_cpu_baseline(J, betas, gammas, cpu_sample_steps)
    return line


# ---- C3: SQA-PT with matrix-core field refresh (the headline) ---------------------------------
def run_c3(steps, warmup, cpu=True):
    """SQA-PT with matrix-core field refresh: dense SK N=2048, P=64, 32 rungs x 4 ladders, 10 sweeps
    per PT step, swaps every PT step, FP32 fields refreshed every 100 sweeps (bf16 hi+lo S.J, FP32
    accumulate) with the FP64 drift check. The configured anneal is 500 PT steps (5000 sweeps); a
    bench step is one twentieth of it, 25 PT steps (250 sweeps of all 128 replicas, 4.19e9 attempts),
    and the K timed steps run as one anneal of 25 K PT steps (K = 20: exactly the 5000-sweep anneal),
    the W warmup steps as one anneal of 25 W PT steps before it. The ladder's (beta, Gamma) points do
    not change during the anneal, so every PT step is the same work. Algorithmic bytes per sweep (128
    replicas): fields r+w 134.2 MB + spins r+w 33.6 MB + J once 16.8 MB = 11.0 B/attempt. One sweep
    launch = 10 sweeps of all 128 replicas (one PT step); its average duration comes from hipEvent
    pairs the library records around each sweep launch on its stream (sweep_ms); the run's device
    time adds the exchange, refresh and drift kernels (3 %)."""
    import torch
    from qlatanneal_amd import abi
    n, P, rungs, L, spb, per_step = 2048, 64, 32, 4, 10, 25
    J = sk_instance(n, 20261204)
    betas = np.geomspace(0.3, 5.0, rungs)
    gammas = np.geomspace(3.0, 0.05, rungs)
    ctx = abi.Context(0)
    dp = abi.DeviceProblem.dense(ctx, np.zeros(n), J)
    attempts = L * rungs * P * n * spb * per_step  # per bench step
    if warmup > 0:
        abi.sqapt_run(dp, betas, gammas, P, spb, 0, per_step * warmup, 1, 0, L, abi.MODE_FP32, 10,
                      refresh_interval=100)
    nsteps = per_step * steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = abi.sqapt_run(dp, betas, gammas, P, spb, 0, nsteps, 1, 0, L, abi.MODE_FP32, 1000, refresh_interval=100)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    best = float(np.min(r["best_energy"]))
    log(f"c3: {nsteps} PT steps, device {r['kernel_ms']:.1f} ms, wall {wall:.2f} s, best {best:.4f}, "
        f"drift {r['field_drift'].tolist()}")
    bytes_per_launch = (2 * L * rungs * P * n * 4 + 2 * L * rungs * P * n + n * n * 4) * spb
    launch_s = float(r["sweep_ms"]) / 1e3 / nsteps  # hipEvent pairs around each sweep launch
    line = _line(steps * attempts / wall, steps, warmup, wall,
                 {"workload": "sqapt_dense_sk_n2048_p64_32rungs_x4_5000sweeps_refresh100_fp32", "N": n,
                  "trotter_slices": P, "rungs": rungs, "ladders": L, "sweeps_per_pt_step": spb,
                  "pt_steps_per_bench_step": per_step, "pt_steps_timed": nsteps,
                  "anneal_pt_steps_configured": 500, "attempts_per_step": attempts,
                  "device_ms": float(r["kernel_ms"]), "sweep_launch_ms_mean": launch_s * 1e3,
                  "field_drift_before_after_refresh": r["field_drift"].tolist(),
                  "best_energy_per_spin": best / n, "accept_rate": float(r["accept_rate"])},
                 _roof(bytes_per_launch, launch_s, traffic_from_profile("sqa_la_kernel_c3_traffic.md"),
                       lockstep=spb * (n + P - 1)))
    if cpu:
        from oracle.orc import Problem
        cores = host_cores()
        lib, kind = cpu_lib(cores)
        p = Problem.dense(np.zeros(n), J)
        t = time.perf_counter()
        lib.sqapt_run(p, betas, gammas, P, spb, 0, 1, 1, 0, 0, 7)  # one ladder, one step
        dt = time.perf_counter() - t
        att = rungs * P * n * spb
        line["cpu_baseline"] = {"value": att / dt, "unit": "attempts/s", "cores": cores, "kind": kind,
                                "sample": (f"one ladder (32 rungs), one PT step (10 sweeps) of the same "
                                           f"instance ({dt:.1f} s); the reference's cost per step does not depend "
                                           f"on the acceptance rate (O(N) delta_energy per attempt)")}
    return line


# ---- C4: the batched-QUBO mix through the per-instance path -----------------------------------
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


def run_c4(steps, warmup, cpu=True, instances=20, streams=4):
    """C4's instances through the plain per-instance GPU path (the batched launch is switched off):
    SQA P=16, 64 replicas per instance (the 64 reads of an instance run together), auto_schedule_sqa
    (50 steps x 20 slice sweeps + 5 worldline sweeps). One timed step = a seeded sample of
    `instances` instances of the C4 mix, driven through the SQAAnnealer class by `streams` host
    threads, largest instance first (each run leases one of the library's HIP contexts, one stream
    each); algorithmic bytes ~10 B per attempt plus J once per instance sweep."""
    import qlatanneal_amd as qa
    P, R, spb, wl = 16, 64, 20, 5
    sched = qa.auto_schedule_sqa()
    nb = len(sched.betas)
    inst = c4_sample(instances)
    attempts = sum(nb * R * (spb * P * n + wl * n) for _, n, _ in inst)
    nbytes = sum(nb * (spb + wl) * (R * P * n * 10 + n * n * 4) for _, n, _ in inst)
    anns = [qa.SQAAnnealer(h.ham if kind == "hubo" else h, sched, P, R, backend="hip") for kind, _, h in inst]
    K = max(1, streams)
    order = sorted(range(len(inst)), key=lambda i: -inst[i][1])

    def one(i, seed0):
        anns[i].set_seed(seed0 + i)
        return anns[i].run(spb, wl).best_energy, anns[i].last_kernel_ms()

    kms, best = [], []
    with ThreadPoolExecutor(K) as pool:
        for w in range(warmup):  # every instance once: its first run on a context uploads the problem there
            list(pool.map(lambda i: one(i, 10 + w), order))
        t0 = time.perf_counter()
        for k in range(steps):
            res = list(pool.map(lambda i: one(i, 1000 + 97 * k), order))
            best.extend(e for e, _ in res)
            kms.append(sum(m for _, m in res))
            log(f"c4 step {k}: {K} thread(s), sweep kernels {kms[-1]:.1f} ms summed over {len(anns)} instances")
        wall = time.perf_counter() - t0
    kern_s = float(np.mean(kms)) / 1e3 if K == 1 else wall / steps
    sizes = sorted(n for _, n, _ in inst)
    line = _line(steps * attempts / wall, steps, warmup, wall,
                 {"workload": f"c4_mix_sample{len(inst)}_sqa_p16_r64_auto50x(20+5)_fp32_per_instance",
                  "instances": len(inst), "hubo": sum(1 for k, _, _ in inst if k == "hubo"),
                  "n_min": sizes[0], "n_median": sizes[len(sizes) // 2], "n_max": sizes[-1],
                  "attempts_per_step": attempts, "sweep_kernel_ms_mean": kern_s * 1e3, "streams": K,
                  "path": ("SQAAnnealer per instance (the batched launch is switched off)"
                           + ("" if K == 1 else f", {K} host threads over the library's HIP streams"))},
                 _roof(nbytes, kern_s),
                 data="synthetic (seeded C4 mix: integer QUBOs of random density and 3-body HUBOs)")
    if cpu:
        from oracle.orc import Problem
        lib, kind = cpu_lib(1)  # one read of one replica: the reference's OpenMP fork is over replicas
        _, n, h = next(x for x in inst if x[0] == "qubo")
        p = Problem.dense(np.asarray(h.fields()), np.asarray(h.couplings()), h.constant())
        nreads = 4
        t = time.perf_counter()
        for r in range(nreads):  # the reference runs an instance's reads one after another
            lib.sqa_run(p, np.array(sched.betas), np.array(sched.gammas), P, 1, spb, wl, 0, 0, 3 + r)
        dt = time.perf_counter() - t
        att = nreads * nb * (spb * P * n + wl * n)
        line["cpu_baseline"] = {"value": att / dt, "unit": "attempts/s", "cores": 1, "kind": kind,
                                "sample": (f"{nreads} of the 64 reads (1 replica each, the full schedule) of the "
                                           f"first QUBO instance (N={n}), {dt:.1f} s")}
    return line


# ---- C5: sqa_chi ------------------------------------------------------------------------------
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


def run_c5(steps, warmup, cpu=True):
    """sqa_chi at the C5 shape with the pilot scan (the method the surrogate replaces): run_chi on a
    dense SK N=2048 and a 3-regular MAX-CUT N=4096, P=32, R=4, beta=4, Gamma 3.0 -> 1e-3, pilot of
    64 points x (10 burn + 30 measured) sweeps, production 100 steps x 10 sweeps. One timed step =
    both runs. Algorithmic bytes ~10 B per attempt; the sweeps are latency-bound."""
    from qlatanneal_amd import abi
    P, R, beta, pts, burn, meas, nsteps, spb = 32, 4, 4.0, 64, 10, 30, 100, 10
    ctx = abi.Context(0)
    n_sk, n_mc = 2048, 4096
    J = sk_instance(n_sk, 20261208)
    probs = [("sk", n_sk, abi.DeviceProblem.dense(ctx, np.zeros(n_sk), J)),
             ("maxcut3", n_mc, abi.DeviceProblem.sparse(ctx, np.zeros(n_mc), maxcut3_edges(n_mc, 20261209), 0.0))]
    kw = dict(num_steps=nsteps, sweeps_per_step=spb, scan_points=pts, scan_sweeps=meas, scan_burn=burn)
    attempts = sum(R * P * n * (pts * (burn + meas) + nsteps * spb) for _, n, _ in probs)
    for w in range(warmup):
        for _, _, dp in probs:
            abi.sqa_chi_run(dp, P, R, beta, 3.0, 1e-3, num_steps=4, sweeps_per_step=2, scan_points=4, scan_sweeps=2,
                            scan_burn=1, seed=10 + w)
    kms, best = [], {}
    t0 = time.perf_counter()
    for k in range(steps):
        tot = 0.0
        for name, n, dp in probs:
            r = abi.sqa_chi_run(dp, P, R, beta, 3.0, 1e-3, seed=1000 + 97 * k, **kw)
            tot += r["kernel_ms"]
            best.setdefault(name, []).append(r["best_energy"] / n)
        kms.append(tot)
        log(f"c5 step {k}: run_chi device {tot:.1f} ms")
    wall = time.perf_counter() - t0
    kern_s = float(np.mean(kms)) / 1e3
    line = _line(steps * attempts / wall, steps, warmup, wall,
                 {"workload": "sqa_chi_run_sk_n2048_and_maxcut3_n4096_p32_r4_pilot64x40_prod100x10_fp32",
                  "trotter_slices": P, "replicas": R, "beta": beta, "scan_points": pts,
                  "pilot_sweeps": burn + meas, "production": [nsteps, spb], "attempts_per_step": attempts,
                  "device_ms_mean": float(np.mean(kms)),
                  "best_energy_per_spin": {k: float(np.min(v)) for k, v in best.items()}},
                 _roof(attempts * 10, kern_s),
                 data="synthetic (seeded SK J_ij ~ N(0, 1/N); 3-regular unit-weight MAX-CUT)")
    if cpu:
        from oracle.orc import Problem
        cores = host_cores()
        lib, kind = cpu_lib(cores)
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
    return line


# ---- time to target ---------------------------------------------------------------------------
def tts99(t_read: float, success: float):
    """mean time per read x ln(0.01) / ln(1 - success rate)
    (examples/python/dwave_comparative_benchmark.py:438-443); None when no read succeeded."""
    if success <= 0.0:
        return None
    if success >= 1.0:
        return t_read
    return t_read * math.log(0.01) / math.log(1.0 - success)


def run_tts(instances=2, reads_a=4, reads_b=8, ladders_pt=32, rungs_pt=4, pt_gamma_max=1.3, gap=0.01, beta_b=16.0,
            workers=4, seed=20261210):
    """Time to reach, with 99 % probability, an energy within `gap` of the best FP64-verified energy
    any method found, on dense SK N=2048 instances (J_ij ~ N(0, 1/N)). Every method runs on the
    same `workers` host threads, each set up before any timing (its HIP context, device copies of
    the instances and of the surrogate, A's annealers, a first launch of every kernel); the FP64
    check of the returned energies runs after each timed region.

    A (the drop-in, plain per-instance GPU SQA): solve(method="sqa")'s defaults -- P=32, one replica
      per read, auto_schedule_sqa (50 steps, beta 0.1 -> 4, Gamma 5 -> 0.01) x (20 slice + 5
      worldline sweeps) -- each read one SQAAnnealer run; reads run on `workers` host threads over
      the library's HIP contexts, as solve() runs seeded reads.
    B (the chosen patterns): FP32 sqa_chi whose Gamma schedule comes from the trained chi_B
      surrogate instead of the pilot scan (the pilot runs only when the surrogate declines) -- P=32,
      R=4, beta=16, 100 x 10 production sweeps on the wavefront kernel -- one read per run, runs of
      all instances concurrent on the same host threads; its time includes the surrogate or pilot.
    B_pt (batched FP32 SQA-PT with the chi_B surrogate): `ladders_pt` ladders of `rungs_pt` rungs in
      one launch per instance, every rung annealed on the surrogate's 100-step Gamma schedule at
      beta=16 (qa_chi_schedule, no beta ramp) from its first step with Gamma <= `pt_gamma_max` on
      (the steps above it are paramagnetic: random states already sample them, and they accept the
      most flips), its beta trailing by a geometric factor down to 0.5 (beta 8 -> 16), replica
      exchange between neighbouring rungs after each PT step of 10 sweeps, FP32 fields refreshed on
      the matrix cores every 100 sweeps; one read per ladder, its time includes the schedule.
    C (B's quality reference): the same sqa_chi with the pilot scan, one run per instance.
    Every returned energy is checked against the FP64 energy of its state (host, reference order).
    Success of a read: E <= E* + gap |E*|, E* the best verified energy over A, B, B_pt and C. A
    method with no success reports a lower bound: the TTS99 it would have with one success in
    one more read."""
    import qlatanneal_amd as qa
    from qlatanneal_amd import abi
    from qlatanneal_amd import surrogate as S
    n, P, R = 2048, 32, 4
    ctx = abi.Context(0)
    sur = S.ChiSurrogate.load(ctx, SURROGATE_MODEL) if os.path.exists(SURROGATE_MODEL) else None
    cfg = sur.cfg if sur is not None else S.PilotConfig()
    sched_a = qa.auto_schedule_sqa()
    Js = [sk_instance(n, seed + k) for k in range(instances)]
    hams = [qa.DenseIsing(np.zeros(n), J) for J in Js]
    dps = [abi.DeviceProblem.dense(ctx, np.zeros(n), J) for J in Js]
    local = threading.local()

    def verify(k, state, e):
        e64 = hams[k].energy([int(x) for x in state])
        assert e64 == e, ("returned energy is not the FP64 energy of its state", k, e, e64)
        return e

    chi_kw = dict(num_steps=100, sweeps_per_step=10, scan_points=cfg.scan_points, scan_sweeps=cfg.scan_sweeps,
                  scan_burn=cfg.scan_burn, driver_A0=cfg.driver_A0)
    # the host threads' contexts, created here one after another: each stream then gets its own
    # hardware queue (the runtime deals them out in creation order), so the workers' runs overlap
    tctx = [abi.Context(0) for _ in range(workers)]
    tnext = iter(range(workers))
    tlock = threading.Lock()
    pool = ThreadPoolExecutor(workers)
    gate = threading.Barrier(workers)

    def setup(_):
        """Per host thread, untimed: A's annealers (one per instance), a context with its own
        device copies of the instances and of the surrogate, and a first launch of every kernel the
        timed runs use (a one-step read of each method), so that no timed region creates a
        context, uploads an instance or loads code."""
        try:
            prepare_thread()
        except BaseException:
            gate.abort()  # the other setup jobs stop waiting and raise too
            raise
        gate.wait()  # one setup job per thread

    def prepare_thread():
        with tlock:
            local.ctx = tctx[next(tnext)]
        local.anns = {k: qa.SQAAnnealer(hams[k], sched_a, P, 1, backend="hip") for k in range(instances)}
        local.dps = [abi.DeviceProblem.dense(local.ctx, np.zeros(n), J) for J in Js]
        local.sur = S.ChiSurrogate.load(local.ctx, SURROGATE_MODEL) if sur is not None else None
        mlp = local.sur.mlp if local.sur is not None else None
        warm_a = qa.SQAAnnealer(hams[0], qa.auto_schedule_sqa(steps=1), P, 1, backend="hip")
        warm_a.set_seed(1)
        warm_a.run(1, 1)
        abi.sqa_chi_run(local.dps[0], cfg.trotter_slices, R, beta_b, cfg.gamma_start, cfg.gamma_end, num_steps=2,
                        sweeps_per_step=1, scan_points=cfg.scan_points, scan_sweeps=cfg.scan_sweeps,
                        scan_burn=cfg.scan_burn, driver_A0=cfg.driver_A0, surrogate=mlp, seed=1)
        sch = abi.chi_schedule(local.dps[0], cfg.trotter_slices, R, beta_b, cfg.gamma_start, cfg.gamma_end,
                               num_steps=2, scan_points=cfg.scan_points, scan_sweeps=cfg.scan_sweeps,
                               scan_burn=cfg.scan_burn, driver_A0=cfg.driver_A0, beta_ramp_fraction=0.0,
                               surrogate=mlp, seed=1)
        sb, sg = S.chi_annealed_ladder(sch["beta_schedule"], sch["gamma_schedule"], rungs_pt, 0.5)
        abi.sqapt_run(local.dps[0], sb[-1], sg[-1], P, 1, 0, 2, 1, 0, ladders=1, seed=1, step_betas=sb,
                      step_gammas=sg, refresh_interval=1)

    list(pool.map(setup, range(workers)))

    def timed_pool(fn, jobs):
        t0 = time.perf_counter()
        out = list(pool.map(fn, jobs))
        return out, time.perf_counter() - t0

    # A: one annealer per (instance, host thread), reads as solve() runs them
    def read_a(job):
        k, r = job
        a = local.anns[k]
        a.set_seed(seed * 1000 + k * 100 + r)
        res = a.run(20, 5)
        return k, res.best_state.spins, res.best_energy

    ra, wall_a = timed_pool(read_a, [(k, r) for k in range(instances) for r in range(reads_a)])
    ea = [(k, verify(k, st, e)) for k, st, e in ra]
    t_read_a = wall_a / (instances * reads_a)
    log(f"tts A: {instances * reads_a} reads in {wall_a:.1f} s")

    # B: sqa_chi on the surrogate schedule, on the thread's context
    def read_b(job):
        k, r = job
        res = abi.sqa_chi_run(local.dps[k], cfg.trotter_slices, R, beta_b, cfg.gamma_start, cfg.gamma_end,
                              surrogate=(local.sur.mlp if local.sur is not None else None),
                              seed=seed + 1000 * k + r, **chi_kw)
        return k, res["best_state"], res["best_energy"], bool(res["surrogate_used"])

    rb, wall_b = timed_pool(read_b, [(k, r) for k in range(instances) for r in range(reads_b)])
    eb = [(k, verify(k, st, e), u) for k, st, e, u in rb]
    t_read_b = wall_b / (instances * reads_b)
    log(f"tts B: {instances * reads_b} runs in {wall_b:.1f} s")

    # B_pt: batched FP32 SQA-PT, every ladder annealed on the chi_B surrogate's schedule
    def ladder_pt(k):
        sch = abi.chi_schedule(local.dps[k], cfg.trotter_slices, R, beta_b, cfg.gamma_start, cfg.gamma_end,
                               num_steps=100, scan_points=cfg.scan_points, scan_sweeps=cfg.scan_sweeps,
                               scan_burn=cfg.scan_burn, driver_A0=cfg.driver_A0, beta_ramp_fraction=0.0,
                               surrogate=(local.sur.mlp if local.sur is not None else None), seed=seed + 7 * k)
        sb, sg = S.chi_annealed_ladder(sch["beta_schedule"], sch["gamma_schedule"], rungs_pt, 0.5,
                                       gamma_max=pt_gamma_max)
        r = abi.sqapt_run(local.dps[k], sb[-1], sg[-1], P, 10, 0, sb.shape[0], 1, 0, ladders=ladders_pt,
                          seed=seed + 13 * k, step_betas=sb, step_gammas=sg, refresh_interval=100)
        return k, r["best_state"], r["best_energy"], bool(sch["surrogate_used"]), sb.shape[0]

    rp, wall_p = timed_pool(ladder_pt, list(range(instances)))
    pool.shutdown()
    ep = [(k, [verify(k, st[l], e[l]) for l in range(ladders_pt)], u) for k, st, e, u, _ in rp]
    pt_steps = [m for *_, m in rp]
    t_read_p = wall_p / (instances * ladders_pt)
    log(f"tts B_pt: {instances * ladders_pt} ladders in {wall_p:.1f} s")

    # C: pilot-scheduled sqa_chi
    ec = []
    for k in range(instances):
        r = abi.sqa_chi_run(dps[k], cfg.trotter_slices, R, beta_b, cfg.gamma_start, cfg.gamma_end,
                            seed=seed + 31 * k, **chi_kw)
        ec.append(verify(k, r["best_state"], r["best_energy"]))

    def tts_or_bound(t_read, success, reads):
        """(TTS99, is a lower bound): with no success in `reads` reads, the TTS99 the method would
        have with one success in one more read -- its true TTS99 is at least that."""
        v = tts99(t_read, success)
        return (v, False) if v is not None else (tts99(t_read, 1.0 / (reads + 1)), True)

    per = []
    for k in range(instances):
        a = np.array([e for kk, e in ea if kk == k])
        b = np.array([e for kk, e, _ in eb if kk == k])
        pt = np.array(next(x[1] for x in ep if x[0] == k))
        pt_sur = next(x[2] for x in ep if x[0] == k)
        best = min(a.min(), b.min(), pt.min(), ec[k])
        thr = best + gap * abs(best)
        sa, sb, sp = (float(np.mean(v <= thr)) for v in (a, b, pt))
        ta, a_bound = tts_or_bound(t_read_a, sa, len(a))
        tb, b_bound = tts_or_bound(t_read_b, sb, len(b))
        tp, p_bound = tts_or_bound(t_read_p, sp, len(pt))
        per.append({"E_best_per_spin": best / n, "success_A": sa, "success_B": sb, "success_B_pt": sp,
                    "tts99_A_s": ta, "tts99_A_is_lower_bound": a_bound, "tts99_B_s": tb,
                    "tts99_B_is_lower_bound": b_bound, "tts99_B_pt_s": tp, "tts99_B_pt_is_lower_bound": p_bound,
                    "best_A_per_spin": float(a.min()) / n, "mean_A_per_spin": float(a.mean()) / n,
                    "best_B_per_spin": float(b.min()) / n, "mean_B_per_spin": float(b.mean()) / n,
                    "best_B_pt_per_spin": float(pt.min()) / n, "mean_B_pt_per_spin": float(pt.mean()) / n,
                    "pilot_sqa_chi_per_spin": ec[k] / n,
                    "gap_B_mean_vs_pilot_sqa_chi": float((b.mean() - ec[k]) / abs(ec[k])),
                    "gap_B_pt_mean_vs_pilot_sqa_chi": float((pt.mean() - ec[k]) / abs(ec[k])),
                    "surrogate_used_B_pt": pt_sur,
                    "surrogate_used_fraction": float(np.mean([u for kk, _, u in eb if kk == k]))})
    med = lambda key: float(np.median([p[key] for p in per]))

    def speedup_kind(m):
        """exact when neither TTS99 is a bound; a lower bound when only A's is (A's true TTS99 is
        larger); indicative when the method's own TTS99 is a bound."""
        if any(p[f"tts99_{m}_is_lower_bound"] for p in per):
            return "indicative"
        return "lower bound" if any(p["tts99_A_is_lower_bound"] for p in per) else "exact"

    return {
        "definition": ("TTS99 = mean wall time per read x ln(0.01)/ln(1 - success rate) "
                       "(examples/python/dwave_comparative_benchmark.py:438-443); success: a read's FP64-verified "
                       f"best energy within {gap:.0%} of the best energy any method found on the instance"),
        "instances": f"{instances} dense SK N=2048, J_ij ~ N(0, 1/N)",
        "A": (f"plain per-instance GPU SQA (solve defaults: P=32, 1 replica/read, 50 x (20+5) sweeps), "
              f"{reads_a} reads/instance on {workers} concurrent HIP contexts"),
        "B": (f"FP32 sqa_chi on the chi_B surrogate's schedule ("
              f"{'model ' + os.path.relpath(SURROGATE_MODEL, ROOT) if sur is not None else 'no model file: pilot scan'}"
              f"), P=32, R=4, beta={beta_b}, 100 x 10 sweeps, {reads_b} runs/instance on {workers} contexts"),
        "B_pt": (f"batched FP32 SQA-PT annealed on the chi_B surrogate's Gamma schedule from Gamma <= "
                 f"{pt_gamma_max} on: {ladders_pt} ladders x {rungs_pt} rungs (beta 8 -> 16) per launch, P=32, "
                 f"{'/'.join(str(m) for m in pt_steps)} PT steps x 10 sweeps, exchange every step, MFMA field "
                 f"refresh every 100 sweeps"),
        "t_read_A_s": t_read_a, "t_read_B_s": t_read_b, "t_read_B_pt_s": t_read_p,
        "tts99_A_s_median": med("tts99_A_s"), "tts99_B_s_median": med("tts99_B_s"),
        "tts99_B_pt_s_median": med("tts99_B_pt_s"),
        "speedup_A_over_B": med("tts99_A_s") / med("tts99_B_s"),
        "speedup_is": speedup_kind("B"),
        "speedup_A_over_B_pt": med("tts99_A_s") / med("tts99_B_pt_s"),
        "speedup_B_pt_is": speedup_kind("B_pt"),
        "gap_B_mean_vs_pilot_sqa_chi": float(np.mean([p["gap_B_mean_vs_pilot_sqa_chi"] for p in per])),
        "gap_B_pt_mean_vs_pilot_sqa_chi": float(np.mean([p["gap_B_pt_mean_vs_pilot_sqa_chi"] for p in per])),
        "per_instance": per,
    }


def _summary(line):
    """The fields of a workload line kept in the headline's other_workloads entry."""
    keep = {k: line[k] for k in ("value", "unit", "ms_per_step", "steps")}
    keep["workload"] = line["config"]["workload"]
    keep["roofline_frac"] = line["roofline"]["frac"]
    if "critical_path" in line["roofline"]:
        keep["ns_per_lockstep_step"] = line["roofline"]["critical_path"]["ns_per_step"]
    return keep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (headline C3: 25 PT steps each, default 20 = the whole 5000-sweep "
                         "anneal; other workloads: whole runs, default 2)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before them (C3: 2, others: 1)")
    ap.add_argument("--cpu-sample-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="default run: only the headline line (no tts99 block, no other workloads)")
    ap.add_argument("--workload", choices=("c1", "c2", "c2v", "c3", "c4", "c5", "tts"), default=None,
                    help="run one workload alone (default: the headline C3 line with tts99 and the others)")
    ap.add_argument("--c4-instances", type=int, default=20)
    ap.add_argument("--c4-streams", type=int, default=4,
                    help="c4: host threads driving the instances (each annealer runs on one of the "
                         "library's HIP streams; 1 runs the instances one after another)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a device copy of known size (torch elementwise multiply by 1 of 256 MiB, "
                         "4 B read + 4 B written per element, 3 times): the FETCH_SIZE / WRITE_SIZE "
                         "calibration pass for the counter-measured HBM traffic")
    ap.add_argument("--tts-instances", type=int, default=2)
    ap.add_argument("--tts-reads", type=int, default=4, help="tts: plain-SQA reads per instance")
    ap.add_argument("--tts-runs", type=int, default=8, help="tts: surrogate-scheduled sqa_chi runs per instance")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every run here has a single MI355X, and the sweep shards by "
              "independent replicas and instances, one GPU each; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_dispatch": a.numel() * 4,
                          "bytes_written_per_dispatch": b.numel() * 4, "dispatches": 3}), flush=True)
        return
    cpu = not args.no_cpu
    w = args.workload
    headline = w in (None, "c3")
    if args.steps is None:
        args.steps = 20 if headline else 2
    if args.warmup is None:
        args.warmup = 2 if headline else 1
    if w == "c1":
        line = run_c1(args.steps, args.warmup, cpu)
    elif w == "c2":
        line = run_c2(args.steps, args.warmup, cpu, args.cpu_sample_steps)
    elif w == "c2v":
        line = run_c2v(args.steps, args.warmup, cpu, cpu_sample_steps=args.cpu_sample_steps)
    elif w == "c3":
        line = run_c3(args.steps, args.warmup, cpu)
    elif w == "c4":
        line = run_c4(args.steps, args.warmup, cpu, args.c4_instances, args.c4_streams)
    elif w == "c5":
        line = run_c5(args.steps, args.warmup, cpu)
    elif w == "tts":
        print(json.dumps(run_tts(args.tts_instances, args.tts_reads, args.tts_runs)), flush=True)
        return
    else:
        line = run_c3(args.steps, args.warmup, cpu)
        if not args.no_extras:
            # the C4 sample first: its concurrent runs share the process's hardware queues with
            # every HIP stream still open, and the tts block opens one context per host thread
            c4 = _summary(run_c4(1, 2, False, args.c4_instances, args.c4_streams))
            others = {"c1": _summary(run_c1(2, 1, False)), "c2": _summary(run_c2(2, 1, False)), "c4_sample": c4,
                      "c5": _summary(run_c5(1, 1, False))}
            line["tts99"] = run_tts(args.tts_instances, args.tts_reads, args.tts_runs)
            line["other_workloads"] = others
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
