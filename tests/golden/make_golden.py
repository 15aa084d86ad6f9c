"""Write the golden fixtures in tests/golden/*.json from the reference build.

Run in the container that holds the reference tree, after `make -C oracle`:

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py              (all fixtures but one)
    OMP_NUM_THREADS=1 python tests/golden/make_golden.py sqa_repeat   (sqa_repeat.json)

Inputs are generated here from fixed numpy seeds and stored in the fixtures themselves, so the
tests never depend on a generator. Outputs come from oracle/_ref/libqanneal_ref.so, the
reference's own sources compiled with its Release flags (oracle/Makefile). It runs with one
OpenMP thread: the reference's sqa_chi parity sweep races when the Trotter slice count is odd
(slices 0 and P-1 share a parity and are neighbours), so only the single-thread order is
reproducible. Floats are stored bit-exactly (float64 little-endian, base64).
"""
from __future__ import annotations

import base64
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.orc import Oracle, Problem, REF_SO, DENSE, SPARSE, HUBO  # noqa: E402


def enc(a) -> str:
    return base64.b64encode(np.ascontiguousarray(a, dtype="<f8").tobytes()).decode()


def problem_json(p: Problem) -> dict:
    d = {"kind": ["dense", "sparse", "hubo"][p.kind], "n": p.n, "c": enc([p.c])}
    if p.h is not None:
        d["h"] = enc(p.h)
    if p.J is not None:
        d["J"] = enc(p.J)
    if p.edges is not None:
        d["edges"] = [[int(i), int(j)] for i, j, _ in p.edges]
        d["edge_values"] = enc([v for _, _, v in p.edges])
    if p.terms is not None:
        d["terms"] = [[int(x) for x in v] for v, _ in p.terms]
        d["term_coeffs"] = enc([c for _, c in p.terms])
    return d


def out_json(res: dict) -> dict:
    o = {}
    for k, v in res.items():
        a = np.asarray(v)
        if a.dtype == np.int8:
            o[k] = a.astype(int).tolist()
        else:
            o[k] = enc(np.atleast_1d(a))
    return o


def sk(rng, n, h_scale=0.0, c=0.0):
    J = rng.normal(0.0, 1.0 / np.sqrt(n), (n, n))
    J = np.triu(J, 1)
    J = J + J.T
    h = rng.normal(0.0, h_scale, n) if h_scale > 0 else np.zeros(n)
    return Problem.dense(h, J, c)


def qubo_problem(ref, rng, n):
    Q = rng.integers(-100, 101, size=(n, n)).astype(float)
    Q = np.triu(Q) + np.triu(Q, 1).T
    h, J, c = ref.qubo_to_ising(Q)
    return Problem.dense(h, J, c), Q


def sparse_problem(rng, n):
    edges = [(i, (i + 1) % n, float(rng.normal())) for i in range(n)]
    edges += [(i, (i + n // 2) % n, float(rng.normal())) for i in range(0, n // 2)]
    return Problem.sparse(rng.normal(0, 0.2, n), edges, 0.5)


def hubo_problem(rng, n, k3):
    terms = [((int(i),), float(rng.normal(0, 0.3))) for i in range(n)]
    for _ in range(k3):
        v = rng.choice(n, size=3, replace=False)
        terms.append((tuple(int(x) for x in v), float(rng.normal())))
    terms.append(((0, 1, 1, 2), 0.7))      # parity reduction: s1^2 = 1 -> (0, 2)
    terms.append(((2, 0), -0.7))          # cancels the merged support above
    terms.append(((3, 3), 1.25))          # folds into c
    return Problem.hubo(n, terms, -0.75)


def main():
    ref = Oracle(REF_SO)
    ref.set_threads(1)
    rng = np.random.default_rng(20261203)

    # ---- Hamiltonian contract ------------------------------------------------------------
    ham = []
    # tests/test_dense_ising.cpp:7-30 known answer: h=(1,-1), J01=0.5, all-up -> E=0.5
    p = Problem.dense([1.0, -1.0], [[0.0, 0.5], [0.5, 0.0]])
    ham.append((p, np.array([[1, 1], [-1, 1]], np.int8)))
    for n in (5, 17, 40):
        ham.append((sk(rng, n, 0.4, 1.5), rng.choice([-1, 1], (4, n)).astype(np.int8)))
    qp, _ = qubo_problem(ref, rng, 12)
    ham.append((qp, rng.choice([-1, 1], (4, 12)).astype(np.int8)))
    ham.append((sparse_problem(rng, 20), rng.choice([-1, 1], (4, 20)).astype(np.int8)))
    ham.append((hubo_problem(rng, 14, 30), rng.choice([-1, 1], (4, 14)).astype(np.int8)))
    cases = []
    for p, s in ham:
        flip = int(rng.integers(0, p.n))
        f = ref.local_fields(p, s[0])
        ns = s[0].copy()
        old = int(ns[flip])
        ns[flip] = -old
        cases.append({
            "problem": problem_json(p), "spins": s.astype(int).tolist(), "flip": flip,
            "energy": enc(ref.energy(p, s)), "delta": enc([ref.delta_energy(p, s[0], flip)]),
            "fields": enc(f), "updated_fields": enc(ref.update_fields(p, f, ns, flip, old)),
        })
    q_cases = []
    for n in (3, 9):
        Q = rng.integers(-100, 101, size=(n, n)).astype(float)
        h, J, c = ref.qubo_to_ising(Q)
        q_cases.append({"Q": enc(Q), "n": n, "h": enc(h), "J": enc(J), "c": enc([c])})
    dump("hamiltonian.json", {"cases": cases, "qubo": q_cases})

    # ---- SA (Annealer::run) --------------------------------------------------------------
    sa = []
    betas_geo = 0.1 * (30.0 ** (np.arange(100) / 99.0))  # geometric 0.1 -> 3.0 over 100 steps
    p64 = sk(rng, 64)
    for seed in (0, 1, 2):
        sa.append(dict(problem=p64, betas=betas_geo, sweeps_per_beta=10, seed=seed))
    sa.append(dict(problem=sk(rng, 8, 0.5, 0.25), betas=np.linspace(0.1, 4.0, 30), sweeps_per_beta=3, seed=7))
    sa.append(dict(problem=qubo_problem(ref, rng, 10)[0], betas=np.linspace(0.01, 0.2, 20), sweeps_per_beta=2, seed=11))
    sa.append(dict(problem=sparse_problem(rng, 16), betas=np.linspace(0.1, 3.0, 25), sweeps_per_beta=2, seed=5))
    sa.append(dict(problem=hubo_problem(rng, 12, 20), betas=np.linspace(0.1, 3.0, 25), sweeps_per_beta=2, seed=6))
    out = []
    for c in sa:
        res = ref.sa_run(c["problem"], c["betas"], c["sweeps_per_beta"], c["seed"])
        out.append({"problem": problem_json(c["problem"]), "betas": enc(c["betas"]),
                    "sweeps_per_beta": c["sweeps_per_beta"], "seed": c["seed"], "out": out_json(res)})
    dump("sa.json", {"cases": out})

    # ---- SQA (SQAAnnealer::run) ----------------------------------------------------------
    sqa = []
    # tests/test_sqa_basic.cpp:8-31 instance and call
    p3 = Problem.dense([0.0] * 3, [[0, .5, -.3], [.5, 0, .2], [-.3, .2, 0]])
    sqa.append(dict(problem=p3, betas=[0.2, 0.5, 1.0], gammas=[2.0, 1.0, 0.5], P=8, R=2, spb=5, wl=1, cl=0, cts=0, seed=42))
    pa = sk(rng, 16, 0.2)
    for (P, R, spb, wl, cl, cts, seed) in [(4, 1, 2, 0, 0, 0, 1), (8, 1, 1, 1, 0, 0, 2), (3, 2, 1, 1, 1, 0, 3),
                                          (1, 1, 2, 1, 0, 0, 4), (2, 3, 1, 0, 1, 0, 5), (4, 2, 1, 0, 0, 6, 6)]:
        sqa.append(dict(problem=pa, betas=np.linspace(0.5, 4.0, 12), gammas=np.geomspace(3.0, 0.01, 12),
                        P=P, R=R, spb=spb, wl=wl, cl=cl, cts=cts, seed=seed))
    p32 = sk(rng, 32)
    sqa.append(dict(problem=p32, betas=np.full(40, 32.0), gammas=np.linspace(3.0, 1e-3, 40), P=8, R=1, spb=1, wl=0, cl=0, cts=0, seed=9))
    sqa.append(dict(problem=qubo_problem(ref, rng, 12)[0], betas=np.linspace(0.01, 0.3, 10), gammas=np.geomspace(50, 1, 10), P=4, R=2, spb=2, wl=1, cl=1, cts=0, seed=10))
    sqa.append(dict(problem=sparse_problem(rng, 20), betas=np.linspace(0.5, 4, 10), gammas=np.geomspace(3, 0.01, 10), P=4, R=2, spb=2, wl=1, cl=1, cts=0, seed=12))
    sqa.append(dict(problem=hubo_problem(rng, 12, 24), betas=np.linspace(0.5, 4, 10), gammas=np.geomspace(3, 0.01, 10), P=4, R=2, spb=2, wl=1, cl=1, cts=0, seed=13))
    out = []
    for c in sqa:
        res = ref.sqa_run(c["problem"], c["betas"], c["gammas"], c["P"], c["R"], c["spb"], c["wl"], c["cl"], c["cts"], c["seed"])
        out.append({"problem": problem_json(c["problem"]), "betas": enc(c["betas"]), "gammas": enc(c["gammas"]),
                    "trotter_slices": c["P"], "replicas": c["R"], "sweeps_per_beta": c["spb"],
                    "worldline_sweeps": c["wl"], "cluster_sweeps": c["cl"], "continuous_time_slices": c["cts"],
                    "seed": c["seed"], "out": out_json(res)})
    dump("sqa.json", {"cases": out})

    # ---- SQA-PT (SQAParallelTemperingAnnealer::run) --------------------------------------
    pt = []
    # tests/test_sqa_parallel_tempering.cpp:7-32 instance and call
    p3b = Problem.dense([0.0] * 3, [[0, .5, -.2], [.5, 0, .3], [-.2, .3, 0]])
    pt.append(dict(problem=p3b, betas=[0.2, 0.5, 1.0], gammas=[2.0, 1.0, 0.5], P=8, spb=4, wl=1, steps=5, si=1, cl=1, cts=0, seed=123))
    pb = sk(rng, 14, 0.2)
    for (P, spb, wl, steps, si, cl, seed) in [(4, 2, 0, 8, 1, 0, 1), (3, 1, 1, 6, 2, 1, 2), (6, 1, 1, 5, 1, 0, 3)]:
        pt.append(dict(problem=pb, betas=np.geomspace(0.3, 3.0, 5), gammas=np.geomspace(3.0, 0.1, 5),
                       P=P, spb=spb, wl=wl, steps=steps, si=si, cl=cl, cts=0, seed=seed))
    out = []
    for c in pt:
        res = ref.sqapt_run(c["problem"], c["betas"], c["gammas"], c["P"], c["spb"], c["wl"], c["steps"], c["si"], c["cl"], c["cts"], c["seed"])
        out.append({"problem": problem_json(c["problem"]), "betas": enc(c["betas"]), "gammas": enc(c["gammas"]),
                    "trotter_slices": c["P"], "sweeps_per_step": c["spb"], "worldline_sweeps": c["wl"],
                    "steps": c["steps"], "swap_interval": c["si"], "cluster_sweeps": c["cl"],
                    "continuous_time_slices": c["cts"], "seed": c["seed"], "out": out_json(res)})
    dump("sqapt.json", {"cases": out})

    # ---- sqa_chi (SQAChiAnnealer::run_chi) -----------------------------------------------
    chi = []
    # tests/test_sqa_chi_annealer.cpp:167-239 (A): SK-like n=10, 8 slices, 4 replicas
    chi.append(dict(problem=sk(rng, 10), P=8, R=4, beta=4.0, gs=4.0, ge=0.05, ns=30, sps=4, sp=12, ss=12, sb=4, seed=2026))
    # (B): dense ferromagnetic chain n=12 (exact ground state -(n-1))
    n = 12
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = J[i + 1, i] = -1.0
    chi.append(dict(problem=Problem.dense(np.zeros(n), J), P=8, R=4, beta=4.0, gs=3.0, ge=0.05, ns=100, sps=10, sp=10, ss=20, sb=6, seed=7))
    # (C): sparse ferromagnetic ring n=12 (exact ground state -n)
    chi.append(dict(problem=Problem.sparse(np.zeros(n), [(i, (i + 1) % n, -1.0) for i in range(n)]), P=8, R=4, beta=4.0, gs=3.0, ge=0.05, ns=100, sps=10, sp=10, ss=20, sb=6, seed=11))
    # (D): degenerate / tiny range
    n = 6
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = J[i + 1, i] = -1.0
    chi.append(dict(problem=Problem.dense(np.zeros(n), J), P=4, R=2, beta=0.5, gs=0.2, ge=0.15, ns=20, sps=4, sp=4, ss=4, sb=2, seed=3))
    # odd slice count (single-thread order) and a HUBO instance
    chi.append(dict(problem=sk(rng, 9), P=5, R=2, beta=2.0, gs=3.0, ge=0.05, ns=12, sps=2, sp=6, ss=4, sb=2, seed=21))
    chi.append(dict(problem=hubo_problem(rng, 10, 15), P=4, R=2, beta=2.0, gs=3.0, ge=0.05, ns=12, sps=2, sp=6, ss=4, sb=2, seed=22))
    out = []
    for c in chi:
        res = ref.sqa_chi_run(c["problem"], c["P"], c["R"], c["beta"], c["gs"], c["ge"], c["ns"], c["sps"], c["sp"], c["ss"], c["sb"], seed=c["seed"])
        out.append({"problem": problem_json(c["problem"]), "trotter_slices": c["P"], "replicas": c["R"],
                    "beta": c["beta"], "gamma_start": c["gs"], "gamma_end": c["ge"], "num_steps": c["ns"],
                    "sweeps_per_step": c["sps"], "scan_points": c["sp"], "scan_sweeps": c["ss"],
                    "scan_burn": c["sb"], "seed": c["seed"], "out": out_json(res)})
    dump("sqa_chi.json", {"cases": out})

    # ---- classical ReplicaAnnealer / ParallelTemperingAnnealer (own stream: the fixtures above
    #      do not move) ---------------------------------------------------------------------------
    rng2 = np.random.default_rng(20261204)
    out = []
    for p, betas, sw, R, seed in [(sk(rng2, 24), np.geomspace(0.1, 3.0, 20), 3, 4, 3),
                                  (sparse_problem(rng2, 18), np.linspace(0.1, 2.5, 15), 2, 3, 9),
                                  (hubo_problem(rng2, 12, 18), np.linspace(0.1, 3.0, 12), 2, 5, 4),
                                  (qubo_problem(ref, rng2, 10)[0], np.linspace(0.01, 0.3, 10), 2, 2, 8)]:
        res = ref.replica_run(p, betas, sw, R, seed)
        out.append({"problem": problem_json(p), "betas": enc(betas), "sweeps_per_beta": sw, "replicas": R,
                    "seed": seed, "out": out_json(res)})
    dump("replica.json", {"cases": out})
    out = []
    for p, betas, sps, steps, si, seed in [(sk(rng2, 20), np.geomspace(0.2, 4.0, 6), 2, 25, 1, 5),
                                           (sparse_problem(rng2, 16), np.geomspace(0.3, 3.0, 4), 3, 20, 2, 6),
                                           (hubo_problem(rng2, 10, 14), np.geomspace(0.2, 3.0, 5), 2, 15, 3, 7)]:
        res = ref.pt_run(p, betas, sps, steps, si, seed)
        out.append({"problem": problem_json(p), "betas": enc(betas), "sweeps_per_step": sps, "steps": steps,
                    "swap_interval": si, "seed": seed, "out": out_json(res)})
    dump("pt.json", {"cases": out})

    # ---- run_optimal: adaptive J_perp schedules (own stream) -------------------------------------
    rng3 = np.random.default_rng(20261205)
    out = []
    for p, P, R, beta, j0, j1, eps, steps, sps, wl, cl, probes, cs, frac, seed in [
            (sk(rng3, 16), 6, 2, 3.0, 0.05, 2.5, 0.0, 14, 2, 1, 0, 5, 2, 0.3, 11),
            (sk(rng3, 12), 4, 1, 2.0, 0.1, 2.0, 0.0, 10, 1, 0, 1, 4, 3, 0.0, 12),
            (sparse_problem(rng3, 14), 5, 3, 2.5, 0.05, 3.0, 0.04, 12, 1, 1, 1, 4, 2, 0.25, 13),
            (hubo_problem(rng3, 10, 14), 4, 2, 3.0, 0.02, 2.0, 0.0, 12, 2, 0, 0, 6, 2, 0.5, 14)]:
        res = ref.sqa_opt_run(p, P, R, beta, j0, j1, eps, 15 / 14, steps, sps, wl, cl, probes, cs, frac, 0.1, seed)
        out.append({"problem": problem_json(p), "trotter_slices": P, "replicas": R, "beta": beta,
                    "j_perp_start": j0, "j_perp_end": j1, "eps_tilde": eps, "alpha": 15 / 14, "num_steps": steps,
                    "sweeps_per_step": sps, "worldline_sweeps": wl, "cluster_sweeps": cl, "calib_probes": probes,
                    "calib_sweeps": cs, "beta_ramp_fraction": frac, "beta_ramp_start": 0.1, "seed": seed,
                    "out": out_json(res)})
    dump("sqa_opt.json", {"cases": out})
    out = []
    for p, nb, P, steps, sps, wl, cl, si, eps, j1, probes, seed in [
            (sk(rng3, 14), 4, 4, 12, 2, 1, 0, 1, 0.0, 2.5, 4, 21),
            (sparse_problem(rng3, 12), 3, 5, 10, 1, 0, 1, 2, 0.0, 0.0, 5, 22),
            (hubo_problem(rng3, 10, 12), 3, 4, 10, 1, 1, 0, 1, 0.05, 2.0, 3, 23),
            (sk(rng3, 10), 2, 3, 8, 2, 0, 0, 1, 0.0, 0.001, 3, 24)]:
        betas = np.geomspace(0.5, 3.0, nb)
        gammas = np.geomspace(2.0, 0.4, nb)
        res = ref.sqapt_opt_run(p, betas, gammas, P, steps, sps, wl, eps, 15 / 14, j1, cl, si, 0, probes, 2, seed)
        out.append({"problem": problem_json(p), "betas": enc(betas), "gammas": enc(gammas), "trotter_slices": P,
                    "num_steps": steps, "sweeps_per_step": sps, "worldline_sweeps": wl, "cluster_sweeps": cl,
                    "swap_interval": si, "eps_tilde": eps, "alpha": 15 / 14, "j_perp_end": j1,
                    "calib_probes": probes, "calib_sweeps": 2, "seed": seed, "out": out_json(res)})
    dump("sqapt_opt.json", {"cases": out})

    # ---- observers: per-step / per-sweep records (own stream) ------------------------------------
    rng4 = np.random.default_rng(20261206)
    sa_cases, sqa_cases = [], []
    for p, betas, sw, seed in [(sk(rng4, 12), np.linspace(0.2, 2.5, 8), 3, 31),
                               (hubo_problem(rng4, 9, 12), np.linspace(0.1, 2.0, 6), 2, 32)]:
        res = ref.sa_observe_run(p, betas, sw, seed)
        sa_cases.append({"problem": problem_json(p), "betas": enc(betas), "sweeps_per_beta": sw, "seed": seed,
                         "out": out_json(res)})
    for p, P, R, spb, wl, cl, tracing, seed in [(sk(rng4, 10), 4, 2, 2, 1, 1, True, 41),
                                                (sparse_problem(rng4, 9), 3, 2, 1, 1, 0, True, 42),
                                                (sk(rng4, 12), 4, 3, 2, 1, 1, False, 43)]:
        betas, gammas = np.linspace(0.5, 3.0, 5), np.geomspace(2.5, 0.1, 5)
        res = ref.sqa_observe_run(p, betas, gammas, P, R, spb, wl, cl, tracing, seed)
        sqa_cases.append({"problem": problem_json(p), "betas": enc(betas), "gammas": enc(gammas),
                          "trotter_slices": P, "replicas": R, "sweeps_per_beta": spb, "worldline_sweeps": wl,
                          "cluster_sweeps": cl, "tracing": tracing, "seed": seed, "out": out_json(res)})
    dump("observers.json", {"sa": sa_cases, "sqa": sqa_cases})


def main_repeat():
    """sqa_repeat.json: consecutive SQAAnnealer::run calls without reseeding (the engine
    continues from run to run); outputs of the last run."""
    ref = Oracle(REF_SO)
    ref.set_threads(1)
    rng = np.random.default_rng(20261204)
    out = []
    for (n, P, R, spb, wl, cl, runs, seed) in [(16, 4, 3, 2, 1, 1, 2, 8), (20, 8, 2, 1, 1, 0, 3, 21)]:
        p = sk(rng, n, 0.2)
        betas, gammas = np.linspace(0.3, 3, 12), np.geomspace(3, 0.02, 12)
        res = ref.sqa_run_repeat(p, runs, betas, gammas, P, R, spb, wl, cl, 0, seed)
        out.append({"problem": problem_json(p), "runs": runs, "betas": enc(betas), "gammas": enc(gammas),
                    "trotter_slices": P, "replicas": R, "sweeps_per_beta": spb, "worldline_sweeps": wl,
                    "cluster_sweeps": cl, "seed": seed, "out": out_json(res)})
    dump("sqa_repeat.json", {"cases": out})


def dump(name, obj):
    obj["source"] = ("reference sources compiled by oracle/Makefile (Release flags), "
                     "OMP_NUM_THREADS=1; written by tests/golden/make_golden.py")
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, separators=(",", ":"))
    print(name, os.path.getsize(os.path.join(HERE, name)), "bytes")


if __name__ == "__main__":
    if sys.argv[1:] == ["sqa_repeat"]:
        main_repeat()
    else:
        main()
