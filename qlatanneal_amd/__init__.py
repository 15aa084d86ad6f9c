"""qlatanneal_amd — the qanneal annealing-sweep API (sa / sqa / sqapt / sqa_chi) on MI355X (gfx950).

Drop-in for the reference's ``qanneal`` package on this path: the same classes (from the
``_qanneal`` extension over the C++17 classes in include/qanneal) and the same ``solve()``.
Every annealing run executes on the GPU through libqanneal_hip.so; importing fails loudly when
the native parts are not built.
"""
try:
    from ._qanneal import *  # noqa: F401,F403
    from ._qanneal import version_string
except ImportError as exc:  # pragma: no cover - exercised only on unbuilt trees
    raise ImportError("qlatanneal_amd native modules are not built; run __graft_entry__.build() "
                      f"(there is no CPU fallback): {exc}") from exc

from .solver import (  # noqa: F401
    SolveResult,
    auto_ladder_sqa_tuned,
    auto_schedule_sa,
    auto_schedule_sa_tuned,
    auto_schedule_sqa,
    auto_schedule_sqa_tuned,
    j_perp_from_beta_gamma,
    j_rms_from_problem,
    optimal_j_perp_params,
    solve,
)

__version__ = version_string()

__all__ = [
    "__version__", "version_string", "version_major", "version_minor", "version_patch", "hip_context_count",
    "State", "Hamiltonian", "DenseIsing", "SparseEdge", "SparseIsing", "QUBO", "HigherOrderIsing",
    "AnnealSchedule", "Observer", "MetricsObserver", "StateTraceObserver", "AnnealResult", "Annealer", "ReplicaResult", "MultiAnnealResult",
    "ReplicaAnnealer", "ParallelTemperingAnnealer", "ParallelTemperingResult", "SQASchedule", "SQAObserver", "SQASweepPhase",
    "SQAMetricsObserver", "SQAStateTraceObserver", "SQAResult",
    "SQAAnnealer", "SQAChiResult", "SQAChiAnnealer", "SQAParallelTemperingAnnealer",
    "SQAParallelTemperingResult", "magnetization", "overlap", "SolveResult", "solve", "auto_schedule_sa",
    "auto_schedule_sqa", "auto_schedule_sa_tuned", "auto_schedule_sqa_tuned", "auto_ladder_sqa_tuned",
    "j_perp_from_beta_gamma", "optimal_j_perp_params", "j_rms_from_problem",
]
