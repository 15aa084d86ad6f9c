"""solve(): the reference's Python front end (python/qanneal/solver.py) over the HIP backend.

Same function names, keyword arguments, defaults and result type as the reference, with two
differences a caller sees: ``backend`` defaults to ``"hip"`` (``"hip_fp64"`` replays the
reference trajectory in FP64; ``"cpu"`` has no sweep engine here and raises), and
``method="ctpimc"`` is not part of this build.
"""
from __future__ import annotations

import logging
import math
import queue
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Any, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from ._qanneal import (
    AnnealSchedule,
    Annealer,
    DenseIsing,
    HigherOrderIsing,
    QUBO,
    SparseIsing,
    SQAAnnealer,
    SQAChiAnnealer,
    SQAParallelTemperingAnnealer,
    SQASchedule,
    hip_context_count,
)

try:  # optional input formats, as in the reference
    import dimod  # type: ignore
except Exception:  # pragma: no cover
    dimod = None
try:
    import networkx as nx  # type: ignore
except Exception:  # pragma: no cover
    nx = None
try:
    from tqdm import tqdm  # type: ignore
except Exception:  # pragma: no cover
    tqdm = None


@dataclass
class SolveResult:
    method: str
    samples: List[np.ndarray]
    energies: List[float]
    best_sample: np.ndarray
    best_energy: float
    trace: Optional[List[float]]
    var_order: List[Any]
    return_bits: bool


def j_perp_from_beta_gamma(beta: float, gamma: float, trotter_slices: int = 32) -> float:
    """J_perp = 0.5 ln(1 / tanh(beta Gamma / M)), with beta Gamma / M clamped at 1e-12 (solver.py:63-68)."""
    x = max(beta * gamma / trotter_slices, 1e-12)
    return 0.5 * math.log(1.0 / math.tanh(x))


def j_rms_from_problem(problem: Any, n: Optional[int] = None) -> Optional[float]:
    """RMS of the present off-diagonal couplings, or None when they cannot be inspected (solver.py:71-125)."""

    def rms(M):
        M = np.asarray(M, dtype=float)
        if M.ndim != 2 or M.shape[0] != M.shape[1]:
            return None
        off = M[np.triu_indices(M.shape[0], k=1)]
        nz = off[off != 0.0]
        return None if nz.size == 0 else float(math.sqrt(float(np.mean(nz * nz))))

    def rms_list(vals):
        vals = [v for v in vals if v != 0.0]
        return None if not vals else float(math.sqrt(sum(v * v for v in vals) / len(vals)))

    try:
        if hasattr(problem, "couplings") and callable(getattr(problem, "couplings")):
            return rms(problem.couplings())
        if isinstance(problem, np.ndarray):
            return rms(problem)
        if nx is not None and isinstance(problem, nx.Graph):
            return rms_list([float(d.get("weight", 0.0)) for _, _, d in problem.edges(data=True)])
        if isinstance(problem, dict):
            return rms_list([float(v) for (i, j), v in problem.items() if i != j])
        if isinstance(problem, (list, tuple)) and problem and isinstance(problem[0], (list, tuple)):
            return rms_list([float(e[2]) for e in problem if int(e[0]) != int(e[1])])
        if dimod is not None and isinstance(problem, dimod.BinaryQuadraticModel):
            return rms_list([float(b) for b in problem.quadratic.values()])
    except Exception:
        return None
    return None


def _mode_params(mode: str) -> Tuple[float, float, float, float, int, int]:
    # (p_start, p_end, gamma_mult, gamma_final_ratio, sa_step_base, sqa_step_base), solver.py:348-357
    m = mode.lower().strip()
    if m == "fast":
        return 0.85, 0.22, 2.0, 0.03, 30, 40
    if m == "accurate":
        return 0.90, 0.04, 4.5, 0.008, 110, 130
    if m != "balanced":
        raise ValueError("mode must be one of: fast, balanced, accurate")
    return 0.88, 0.10, 3.0, 0.015, 60, 80


def _estimate_delta_scale(ising: Any, probes: int = 256, seed: int = 1234) -> float:
    """75th percentile of |dE| over random single-flip probes (solver.py:360-379)."""
    n = int(ising.size())
    if n <= 0:
        return 1.0
    rng = np.random.default_rng(seed)
    k = int(max(32, min(probes, max(64, n * 4))))
    deltas: List[float] = []
    for _ in range(k):
        spins = rng.choice([-1, 1], size=n).astype(int).tolist()
        flip = int(rng.integers(0, n))
        try:
            d = abs(float(ising.delta_energy(spins, flip)))
            if np.isfinite(d):
                deltas.append(d)
        except Exception:
            continue
    if not deltas:
        return 1.0
    return max(float(np.quantile(np.array(deltas, dtype=float), 0.75)), 1e-6)


def optimal_j_perp_params(problem: Any, mode: str = "balanced", trotter_slices: int = 32, probes: int = 256,
                          seed: int = 1234, n: Optional[int] = None) -> Tuple[float, float, float]:
    """(beta, j_perp_start, j_perp_end) with beta = max(4 / j_rms, 1) (solver.py:128-167)."""
    ising, _, _ = _normalize_problem(problem, n=n)
    _, p_end, gamma_mult, gamma_final_ratio, _, _ = _mode_params(mode)
    scale = _estimate_delta_scale(ising, probes=probes, seed=seed)
    jrms = j_rms_from_problem(problem, n=n)
    if jrms is not None and jrms > 0.0:
        beta_end = max(4.0 / jrms, 1.0)
    else:
        beta_end = -np.log(p_end) / scale
    gamma_start = max(gamma_mult * scale, 1e-3)
    gamma_end = max(gamma_start * gamma_final_ratio, 1e-6)
    return (float(beta_end), j_perp_from_beta_gamma(float(beta_end), float(gamma_start), trotter_slices),
            j_perp_from_beta_gamma(float(beta_end), float(gamma_end), trotter_slices))


def auto_schedule_sa(steps: int = 50, beta_start: float = 0.1, beta_end: float = 4.0) -> AnnealSchedule:
    return AnnealSchedule.linear(beta_start, beta_end, steps)


def auto_schedule_sqa(steps: int = 50, beta_start: float = 0.1, beta_end: float = 4.0, gamma_start: float = 5.0,
                      gamma_end: float = 0.01) -> SQASchedule:
    """Linear beta, geometric Gamma (solver.py:176-186)."""
    return SQASchedule.from_vectors(np.linspace(beta_start, beta_end, steps).tolist(),
                                    np.geomspace(gamma_start, gamma_end, steps).tolist())


def auto_schedule_sa_tuned(problem: Any, mode: str = "balanced", steps: Optional[int] = None, probes: int = 256,
                           seed: int = 1234, n: Optional[int] = None, beta_end_scale: float = 1.0) -> AnnealSchedule:
    ising, _, _ = _normalize_problem(problem, n=n)
    p_start, p_end, _, _, sa_base, _ = _mode_params(mode)
    scale = _estimate_delta_scale(ising, probes=probes, seed=seed)
    beta_start = -np.log(p_start) / scale
    beta_end = (-np.log(p_end) / scale) * max(beta_end_scale, 1e-6)
    if steps is None:
        steps = int(sa_base + 6 * np.sqrt(max(int(ising.size()), 1)))
    return AnnealSchedule.linear(float(beta_start), float(beta_end), max(int(steps), 4))


def auto_schedule_sqa_tuned(problem: Any, mode: str = "balanced", steps: Optional[int] = None, probes: int = 256,
                            seed: int = 1234, n: Optional[int] = None, beta_end_scale: float = 1.0,
                            gamma_end_scale: float = 1.0) -> SQASchedule:
    ising, _, _ = _normalize_problem(problem, n=n)
    p_start, p_end, gamma_mult, gamma_final_ratio, _, sqa_base = _mode_params(mode)
    scale = _estimate_delta_scale(ising, probes=probes, seed=seed)
    beta_start = -np.log(p_start) / scale
    beta_end = (-np.log(p_end) / scale) * max(beta_end_scale, 1e-6)
    gamma_start = max(gamma_mult * scale, 1e-3)
    gamma_end = max(gamma_start * gamma_final_ratio * max(gamma_end_scale, 1e-8), 1e-6)
    if steps is None:
        steps = int(sqa_base + 8 * np.sqrt(max(int(ising.size()), 1)))
    steps = max(int(steps), 6)
    return SQASchedule.from_vectors(np.linspace(beta_start, beta_end, steps).tolist(),
                                    np.geomspace(gamma_start, gamma_end, steps).tolist())


def auto_ladder_sqa_tuned(problem: Any, replicas: int = 8, mode: str = "balanced", probes: int = 256,
                          seed: int = 1234, n: Optional[int] = None) -> SQASchedule:
    """Geometric (beta, Gamma) ladder for SQA-PT (solver.py:433-452)."""
    ising, _, _ = _normalize_problem(problem, n=n)
    _, p_end, gamma_mult, _, _, _ = _mode_params(mode)
    scale = _estimate_delta_scale(ising, probes=probes, seed=seed)
    replicas = max(int(replicas), 2)
    beta_min = max(-np.log(0.75) / scale, 1e-4)
    beta_max = max(-np.log(p_end) / scale, beta_min * 1.5)
    gamma_max = max(gamma_mult * scale, 1e-3)
    gamma_min = max(gamma_max * 0.02, 1e-4)
    return SQASchedule.from_vectors(np.geomspace(beta_min, beta_max, replicas).tolist(),
                                    np.geomspace(gamma_max, gamma_min, replicas).tolist())


def _guess_n_from_entries(entries: Iterable[Tuple[int, int, float]]) -> int:
    return max((max(i, j) for i, j, _ in entries), default=-1) + 1


def _qubo_from_graph(graph: Any):
    if nx is None:
        raise ValueError("networkx not available; install `networkx` to use graph inputs.")
    nodes = list(graph.nodes())
    index = {v: k for k, v in enumerate(nodes)}
    Q = np.zeros((len(nodes), len(nodes)))
    for v, data in graph.nodes(data=True):
        Q[index[v], index[v]] += float(data.get("bias", 0.0))
    for u, v, data in graph.edges(data=True):
        w = float(data.get("weight", 0.0))
        Q[index[u], index[v]] += w
        Q[index[v], index[u]] += w
    return QUBO(Q), nodes


def _is_higher_order_keys(keys: Iterable[Any]) -> bool:
    for k in keys:
        if isinstance(k, (tuple, list, frozenset, set)):
            if len(k) != 2:
                return True
        else:
            return True
    return False


def _hubo_from_dict(problem: dict, n: Optional[int]):
    max_idx = -1
    for k in problem:
        for i in ((k,) if isinstance(k, int) else tuple(k)):
            max_idx = max(max_idx, int(i))
    size = n if n is not None else max_idx + 1
    return HigherOrderIsing(problem, size), list(range(size))


def _hubo_from_binary_polynomial(poly: Any):
    variables = list(poly.variables)
    index = {v: k for k, v in enumerate(variables)}
    is_binary = "BINARY" in str(getattr(poly, "vartype", "SPIN"))
    terms: dict = {}
    for term, bias in poly.items():
        idxs = [index[v] for v in term]
        bias = float(bias)
        if not is_binary:
            key = tuple(sorted(idxs))
            terms[key] = terms.get(key, 0.0) + bias
            continue
        d = len(idxs)  # x = (1 + s) / 2 expanded over all subsets
        scale = bias / (2 ** d) if d > 0 else bias
        for mask in range(1 << d):
            key = tuple(sorted(idxs[b] for b in range(d) if (mask >> b) & 1))
            terms[key] = terms.get(key, 0.0) + scale
    return HigherOrderIsing(terms, len(variables)), variables, is_binary


def _normalize_problem(problem: Any, n: Optional[int] = None):
    """(ising, var_order, is_qubo) for every input format the reference accepts (solver.py:284-335)."""
    if isinstance(problem, (DenseIsing, SparseIsing, HigherOrderIsing)):
        return problem, list(range(problem.size())), False
    if isinstance(problem, QUBO):
        return problem.to_ising(), list(range(problem.size())), True
    if dimod is not None and isinstance(problem, dimod.BinaryQuadraticModel):
        return QUBO(problem).to_ising(), list(problem.variables), True
    if dimod is not None and isinstance(problem, dimod.BinaryPolynomial):
        return _hubo_from_binary_polynomial(problem)
    if nx is not None and isinstance(problem, nx.Graph):
        q, order = _qubo_from_graph(problem)
        return q.to_ising(), order, True
    if isinstance(problem, np.ndarray):
        return QUBO(problem).to_ising(), list(range(problem.shape[0])), True
    if isinstance(problem, dict):
        if _is_higher_order_keys(problem.keys()):
            ham, order = _hubo_from_dict(problem, n)
            return ham, order, False
        if n is None:
            n = _guess_n_from_entries([(i, j, float(v)) for (i, j), v in problem.items()])
        return QUBO(problem, n).to_ising(), list(range(n)), True
    if isinstance(problem, (list, tuple)) and problem and isinstance(problem[0], (list, tuple)):
        entries = [(int(e[0]), int(e[1]), float(e[2])) for e in problem]
        if n is None:
            n = _guess_n_from_entries(entries)
        return QUBO(entries, n).to_ising(), list(range(n)), True
    raise ValueError("Unsupported problem type. Provide QUBO, DenseIsing, SparseIsing, HigherOrderIsing, dimod "
                     "BQM/BinaryPolynomial, networkx graph, ndarray, dict, or entries list.")


def _iter_reads(total: int, enabled: bool, logger: logging.Logger):
    if total <= 1:
        return range(total)
    if enabled and tqdm is not None:
        return tqdm(range(total), total=total, desc="qanneal")
    if enabled:
        logger.info("Running %d reads...", total)
    return range(total)


def solve(problem: Any,
          method: str = "sqa",
          reads: int = 1,
          sweeps_per_beta: int = 20,
          worldline_sweeps: int = 5,
          cluster_sweeps: int = 0,
          continuous_time_slices: int = 0,
          trotter_slices: int = 32,
          replicas: int = 1,
          ctpimc_qubits_per_update: int = 1,
          ctpimc_qubits_per_chain: int = 1,
          pt_steps: int = 50,
          swap_interval: int = 1,
          pt_betas: Optional[Sequence[float]] = None,
          pt_gammas: Optional[Sequence[float]] = None,
          schedule: Optional[Any] = None,
          seed: Optional[int] = None,
          backend: str = "hip",
          progress: bool = True,
          logger: Optional[logging.Logger] = None,
          n: Optional[int] = None,
          return_bits: bool = False,
          schedule_type: str = "standard",
          optimal_eps_tilde: float = 0.0,
          optimal_alpha: float = 15.0 / 14.0,
          optimal_num_steps: Optional[int] = None,
          optimal_j_perp_start: Optional[float] = None,
          optimal_j_perp_end: Optional[float] = None,
          optimal_beta: Optional[float] = None,
          optimal_calib_probes: int = 12,
          optimal_calib_sweeps: int = 10,
          optimal_beta_ramp_fraction: float = 0.3,
          optimal_debug_csv: Optional[str] = None,
          chi_gamma_start: Optional[float] = None,
          chi_gamma_end: Optional[float] = None,
          chi_scan_points: int = 16,
          chi_scan_sweeps: int = 30,
          chi_scan_burn: int = 10,
          chi_floor_fraction: float = 1e-6,
          chi_driver_A0: float = 0.0) -> SolveResult:
    """Solve with "sa", "sqa", "sqapt" or "sqa_chi" (reference solve(), solver.py:455-789).

    Read r of a seeded call is set_seed(seed + r) and one run, as in the reference's read loop, so
    the call is reproducible. Seeded reads run concurrently, on one engine per HIP context
    (hip_context_count()), unless a debug CSV is requested. Seeded SA reads run in one launch
    (Annealer.run_reads): with backend "hip_fp64" read r is then exactly set_seed(seed + r) and one
    run; with the FP32 backend the launch's random stream is keyed by (seed, read index), which is
    reproducible but not the stream a separate set_seed(seed + r) run would draw. Unseeded reads
    run one after another on one engine, continuing its random stream. Every run executes on the GPU.
    """
    method = method.lower().strip()
    schedule_type = schedule_type.lower().strip()
    if method == "ctpimc":
        raise ValueError("method='ctpimc' is not part of this build (methods: 'sa', 'sqa', 'sqapt', 'sqa_chi')")
    if method not in ("sa", "sqa", "sqapt", "sqa_chi"):
        raise ValueError("method must be 'sa', 'sqa', 'sqapt', 'ctpimc', or 'sqa_chi'")
    if schedule_type not in ("standard", "optimal"):
        raise ValueError("schedule_type must be 'standard' or 'optimal'")
    if method == "sqa_chi" and schedule_type != "standard":
        raise ValueError("method='sqa_chi' does not support schedule_type overrides; "
                         "leave schedule_type='standard' and use the chi_* parameters.")
    if logger is None:
        logger = logging.getLogger("qanneal.solve")

    ising, var_order, is_qubo = _normalize_problem(problem, n=n)
    if return_bits and not is_qubo:
        raise ValueError("return_bits=True requires a QUBO-type input (QUBO, ndarray, dict, entries list, "
                         "networkx graph, or dimod BQM). DenseIsing and SparseIsing use Ising spins {-1, +1}; "
                         "set return_bits=False or convert your problem to QUBO first.")

    if schedule is None:
        if method in ("sqa", "sqa_chi"):
            schedule = auto_schedule_sqa()
        elif method == "sqapt":
            schedule = auto_ladder_sqa_tuned(ising, replicas=max(2, replicas))
        else:
            schedule = auto_schedule_sa()

    if method == "sqapt":
        if pt_betas is not None or pt_gammas is not None:
            if pt_betas is None or pt_gammas is None:
                raise ValueError("pt_betas and pt_gammas must both be provided.")
            lb, lg = list(map(float, pt_betas)), list(map(float, pt_gammas))
        elif isinstance(schedule, SQASchedule) and len(schedule.betas) >= 2 and len(schedule.betas) == len(schedule.gammas):
            lb, lg = list(schedule.betas), list(schedule.gammas)
        else:
            ladder = auto_ladder_sqa_tuned(ising, replicas=max(2, replicas))
            lb, lg = ladder.betas, ladder.gammas

    def make_annealer():
        if method == "sqa":
            return SQAAnnealer(ising, schedule, trotter_slices, replicas, backend=backend)
        if method == "sqa_chi":
            return SQAChiAnnealer(ising, trotter_slices, replicas, backend=backend)
        if method == "sqapt":
            return SQAParallelTemperingAnnealer(ising, lb, lg, trotter_slices, backend=backend)
        return Annealer(ising, schedule, backend=backend)

    annealer = make_annealer()

    opt = {}
    if schedule_type == "optimal" and method in ("sqa", "sqapt"):
        b, jps, jpe = optimal_j_perp_params(ising, trotter_slices=trotter_slices)
        opt["beta"] = float(optimal_beta) if optimal_beta is not None else b
        opt["jps"] = float(optimal_j_perp_start) if optimal_j_perp_start is not None else jps
        if optimal_j_perp_end is not None:
            opt["jpe"] = float(optimal_j_perp_end)
        else:
            jrms = j_rms_from_problem(problem, n=n)
            opt["jpe"] = max(opt["jps"], 5.0 * jrms) if (jrms is not None and jrms > 0.0) else jpe
        opt["steps"] = int(optimal_num_steps) if optimal_num_steps is not None else max(
            100, len(getattr(schedule, "betas", [None] * 100)))

    chi = {}
    if method == "sqa_chi":
        b, _, _ = optimal_j_perp_params(ising, trotter_slices=trotter_slices)
        chi["beta"] = float(optimal_beta) if optimal_beta is not None else b
        gs = list(getattr(schedule, "gammas", []) or [])
        chi["gs"] = float(chi_gamma_start) if chi_gamma_start is not None else (float(max(gs)) if gs else 5.0)
        chi["ge"] = float(chi_gamma_end) if chi_gamma_end is not None else (float(min(gs)) if gs else 0.01)
        if not (chi["gs"] > chi["ge"] > 0.0):
            raise ValueError(f"sqa_chi requires gamma_start > gamma_end > 0 (got {chi['gs']} and {chi['ge']}).")
        chi["steps"] = int(optimal_num_steps) if optimal_num_steps is not None else max(100, len(gs) if gs else 100)

    samples: List[np.ndarray] = []
    energies: List[float] = []
    trace: Optional[List[float]] = None
    if method == "sa" and seed is not None and reads > 1:
        # Seeded SA reads are independent chains: run them in one launch (Annealer.run_reads). In
        # validation mode read r replays set_seed(seed + r) exactly, as the loop below would; the
        # FP32 launch keys its Philox stream with (seed, read index) instead.
        batch = annealer.run_reads(sweeps_per_beta, reads, seed)
        for r in _iter_reads(reads, progress, logger):
            res = batch[r]
            spins = np.array(res.best_state.spins, dtype=int)
            samples.append(((spins + 1) // 2).astype(int) if (return_bits and is_qubo) else spins)
            energies.append(float(res.best_energy))
            if trace is None:
                trace = list(res.energy_trace)
        best = int(np.argmin(energies))
        return SolveResult(method=method, samples=samples, energies=energies, best_sample=samples[best],
                           best_energy=energies[best], trace=trace, var_order=var_order, return_bits=return_bits)

    def one_read(ann, r):
        if seed is not None:
            ann.set_seed(seed + int(r))
        if schedule_type == "optimal" and method == "sqa":
            res = ann.run_optimal(beta=opt["beta"], j_perp_start=opt["jps"], j_perp_end=opt["jpe"],
                                  eps_tilde=optimal_eps_tilde, alpha=optimal_alpha, num_steps=opt["steps"],
                                  sweeps_per_step=sweeps_per_beta, worldline_sweeps=worldline_sweeps,
                                  cluster_sweeps=cluster_sweeps, calib_probes=int(optimal_calib_probes),
                                  calib_sweeps=int(optimal_calib_sweeps),
                                  debug_csv_path=(optimal_debug_csv or ""),
                                  beta_ramp_fraction=float(optimal_beta_ramp_fraction))
            local = list(res.energy_trace)
        elif method == "sqa_chi":
            res = ann.run_chi(beta=chi["beta"], gamma_start=chi["gs"], gamma_end=chi["ge"],
                              num_steps=chi["steps"], sweeps_per_step=sweeps_per_beta,
                              scan_points=int(chi_scan_points), scan_sweeps=int(chi_scan_sweeps),
                              scan_burn=int(chi_scan_burn), chi_floor_fraction=float(chi_floor_fraction),
                              driver_A0=float(chi_driver_A0),
                              beta_ramp_fraction=float(optimal_beta_ramp_fraction),
                              debug_csv_path=(optimal_debug_csv or ""))
            local = list(res.energy_trace)
        elif schedule_type == "optimal" and method == "sqapt":
            res = ann.run_optimal(num_steps=opt["steps"], sweeps_per_step=sweeps_per_beta,
                                  worldline_sweeps=worldline_sweeps, eps_tilde=optimal_eps_tilde,
                                  alpha=optimal_alpha, j_perp_end=opt["jpe"], cluster_sweeps=cluster_sweeps,
                                  swap_interval=swap_interval, continuous_time_slices=continuous_time_slices,
                                  calib_probes=int(optimal_calib_probes),
                                  calib_sweeps=int(optimal_calib_sweeps),
                                  debug_csv_path=(optimal_debug_csv or ""))
            local = list(res.average_energy_trace)
        elif method == "sqa":
            res = ann.run(sweeps_per_beta=sweeps_per_beta, worldline_sweeps=worldline_sweeps,
                          cluster_sweeps=cluster_sweeps, continuous_time_slices=continuous_time_slices)
            local = list(res.energy_trace)
        elif method == "sqapt":
            res = ann.run(sweeps_per_step=sweeps_per_beta, worldline_sweeps=worldline_sweeps, steps=pt_steps,
                          swap_interval=swap_interval, cluster_sweeps=cluster_sweeps,
                          continuous_time_slices=continuous_time_slices)
            local = list(res.average_energy_trace)
        else:
            res = ann.run(sweeps_per_beta=sweeps_per_beta)
            local = list(res.energy_trace)
        return res, local

    def collect(res, local):
        nonlocal trace
        spins = np.array(res.best_state.spins, dtype=int)
        samples.append(((spins + 1) // 2).astype(int) if (return_bits and is_qubo) else spins)
        energies.append(float(res.best_energy))
        if trace is None:
            trace = local

    workers = min(reads, hip_context_count()) if (seed is not None and not optimal_debug_csv) else 1
    if workers > 1:
        # Seeded reads are independent: read r is set_seed(seed + r) and one run, whatever ran
        # before on the engine. One engine per HIP context, each driven by its own thread (the
        # bindings release the GIL), runs the reads concurrently on the GPU; results are
        # collected in read order.
        engines = queue.SimpleQueue()
        engines.put(annealer)
        for _ in range(workers - 1):
            engines.put(make_annealer())

        def task(r):
            ann = engines.get()
            try:
                return one_read(ann, r)
            finally:
                engines.put(ann)

        with ThreadPoolExecutor(workers) as ex:
            futures = [ex.submit(task, r) for r in range(reads)]
            try:
                for r in _iter_reads(reads, progress, logger):
                    collect(*futures[r].result())
            except BaseException:
                for f in futures:  # a failed read ends the call without running the queued ones
                    f.cancel()
                raise
    else:
        for r in _iter_reads(reads, progress, logger):
            collect(*one_read(annealer, r))

    best = int(np.argmin(energies))
    return SolveResult(method=method, samples=samples, energies=energies, best_sample=samples[best],
                       best_energy=energies[best], trace=trace, var_order=var_order, return_bits=return_bits)
