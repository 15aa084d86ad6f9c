"""CPU throughput of the restatement (liboracle.so) against the reference build — TEST
INFRASTRUCTURE ONLY, run where the reference tree was built (oracle/_ref present).

The GPU box times liboracle.so as the CPU baseline (the reference build stays with the reference
tree). This script checks, on the same inputs and thread count, that the restatement runs within
+-10 % of the reference build's portable (-march=x86-64-v3) and native builds, for the samples the
benchmark times: SA C1 (32 sequential reads, 1 thread), SQA C2 (N=512, P=32, R=256, first steps),
SQA-PT C3 (N=2048, P=64, one 32-rung ladder, one step of 10 sweeps).

    python oracle/speed_check.py [--threads T]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from oracle.orc import Oracle, Problem, ORACLE_SO, REF_SO, REF_PORTABLE_SO  # noqa: E402


def sk(n, seed):
    rng = np.random.default_rng(seed)
    J = rng.normal(0.0, 1.0 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    return J + J.T


def samples():
    J1 = sk(64, 20261203)
    b1 = 0.1 * (30.0 ** (np.arange(100) / 99))
    p1 = Problem.dense(np.zeros(64), J1)

    def c1(lib):
        for r in range(32):
            lib.sa_run(p1, b1, 10, 1000 + r)
        return 32 * 100 * 10 * 64

    J2 = sk(512, 20261203)
    p2 = Problem.dense(np.zeros(512), J2)
    b2, g2 = np.full(8, 32.0), np.linspace(3.0, 1e-3, 2000)[:8]

    def c2(lib):
        lib.sqa_run(p2, b2, g2, 32, 256, 1, 0, 0, 0, 1)
        return 8 * 256 * 32 * 512

    J3 = sk(2048, 20261204)
    p3 = Problem.dense(np.zeros(2048), J3)
    b3, g3 = np.geomspace(0.3, 5.0, 32), np.geomspace(3.0, 0.05, 32)

    def c3(lib):
        lib.sqapt_run(p3, b3, g3, 64, 10, 0, 1, 1, 0, 0, 7)
        return 32 * 64 * 2048 * 10

    return [("c1_sa_n64_32reads", c1, 1), ("c2_sqa_n512_p32_r256_8steps", c2, None),
            ("c3_sqapt_n2048_p64_1ladder_1step", c3, None)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    libs = [("restatement", ORACLE_SO), ("reference_portable", REF_PORTABLE_SO), ("reference_native", REF_SO)]
    libs = [(k, p) for k, p in libs if os.path.exists(p)]
    print(f"threads {a.threads}; libraries: {', '.join(k for k, _ in libs)}")
    print("| sample | " + " | ".join(f"{k} attempts/s" for k, _ in libs) + " | restatement / reference_portable |")
    print("|---|" + "---|" * (len(libs) + 1))
    for name, fn, thr in samples():
        rates = []
        for _, path in libs:
            lib = Oracle(path)
            lib.set_threads(thr or a.threads)
            fn(lib)  # warm (page-in, thread pool)
            t = time.perf_counter()
            att = fn(lib)
            rates.append(att / (time.perf_counter() - t))
        ratio = rates[0] / rates[1] if len(rates) > 1 else float("nan")
        print(f"| {name} ({thr or a.threads} threads) | " + " | ".join(f"{r:.3e}" for r in rates) + f" | {ratio:.3f} |",
              flush=True)


if __name__ == "__main__":
    main()
