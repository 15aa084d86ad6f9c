"""ctypes binding of oracle/orc_abi.h — TEST INFRASTRUCTURE ONLY.

Loads either my CPU restatement (``oracle/liboracle.so``) or the reference build
(``oracle/_ref/libqanneal_ref*.so``). Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libqanneal_ref.so")
REF_PORTABLE_SO = os.path.join(HERE, "_ref", "libqanneal_ref_portable.so")

DENSE, SPARSE, HUBO = 0, 1, 2

_i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_bp = C.POINTER(C.c_int8)
_lp = C.POINTER(C.c_int64)


class _Problem(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("n", _i64), ("h", _dp), ("J", _dp), ("c", C.c_double),
        ("nedges", _i64), ("ei", _lp), ("ej", _lp), ("ev", _dp),
        ("nterms", _i64), ("term_off", _lp), ("term_vars", _lp), ("term_coeff", _dp),
    ]


def _d(a):
    return a.ctypes.data_as(_dp)


def _b(a):
    return a.ctypes.data_as(_bp)


def _l(a):
    return a.ctypes.data_as(_lp)


@dataclass
class Problem:
    """A dense, sparse or higher-order Ising problem in the oracle's input form."""

    kind: int
    n: int
    h: Optional[np.ndarray] = None
    J: Optional[np.ndarray] = None
    c: float = 0.0
    edges: Optional[Sequence] = None          # sparse: [(i, j, v), ...]
    terms: Optional[Sequence] = None          # hubo: [((i, j, ...), coeff), ...]

    @staticmethod
    def dense(h, J, c=0.0):
        h = np.ascontiguousarray(h, dtype=np.float64)
        J = np.ascontiguousarray(J, dtype=np.float64)
        return Problem(DENSE, int(h.shape[0]), h=h, J=J, c=float(c))

    @staticmethod
    def sparse(h, edges, c=0.0):
        h = np.ascontiguousarray(h, dtype=np.float64)
        return Problem(SPARSE, int(h.shape[0]), h=h, edges=list(edges), c=float(c))

    @staticmethod
    def hubo(n, terms, c=0.0):
        return Problem(HUBO, int(n), terms=list(terms), c=float(c))

    def _struct(self):
        keep = []
        s = _Problem()
        s.kind, s.n, s.c = self.kind, self.n, self.c
        if self.h is not None:
            keep.append(self.h)
            s.h = _d(self.h)
        if self.J is not None:
            keep.append(self.J)
            s.J = _d(self.J)
        if self.edges is not None:
            e = np.array([(int(a), int(b)) for a, b, _ in self.edges], dtype=np.int64).reshape(-1, 2)
            ei, ej = np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1])
            ev = np.array([float(v) for _, _, v in self.edges], dtype=np.float64)
            keep += [ei, ej, ev]
            s.nedges, s.ei, s.ej, s.ev = len(self.edges), _l(ei), _l(ej), _d(ev)
        if self.terms is not None:
            off = [0]
            vars_ = []
            for v, _ in self.terms:
                vars_ += [int(x) for x in v]
                off.append(len(vars_))
            off = np.array(off, dtype=np.int64)
            vv = np.array(vars_ if vars_ else [0], dtype=np.int64)
            cc = np.array([float(c) for _, c in self.terms] or [0.0], dtype=np.float64)
            keep += [off, vv, cc]
            s.nterms, s.term_off, s.term_vars, s.term_coeff = len(self.terms), _l(off), _l(vv), _d(cc)
        return s, keep


class OracleError(RuntimeError):
    pass


class Oracle:
    """One loaded oracle-ABI library (restatement or reference build)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run `make -C oracle`")
        self.path = path
        self.lib = C.CDLL(path)
        self.lib.orc_last_error.restype = C.c_char_p

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(self.lib.orc_last_error().decode())

    def set_threads(self, k: int) -> None:
        self.lib.orc_set_threads(C.c_int(int(k)))

    def energy(self, p: Problem, spins) -> np.ndarray:
        s = np.ascontiguousarray(spins, dtype=np.int8).reshape(-1, p.n)
        out = np.zeros(s.shape[0])
        st, keep = p._struct()
        self._chk(self.lib.orc_energy(C.byref(st), _i64(s.shape[0]), _b(s), _d(out)))
        return out

    def delta_energy(self, p: Problem, spins, flip: int) -> float:
        s = np.ascontiguousarray(spins, dtype=np.int8)
        out = C.c_double()
        st, keep = p._struct()
        self._chk(self.lib.orc_delta_energy(C.byref(st), _b(s), _i64(flip), C.byref(out)))
        return out.value

    def local_fields(self, p: Problem, spins) -> np.ndarray:
        s = np.ascontiguousarray(spins, dtype=np.int8)
        f = np.zeros(p.n)
        st, keep = p._struct()
        self._chk(self.lib.orc_local_fields(C.byref(st), _b(s), _d(f)))
        return f

    def update_fields(self, p: Problem, fields, new_spins, flip: int, old: int) -> np.ndarray:
        f = np.array(fields, dtype=np.float64)
        s = np.ascontiguousarray(new_spins, dtype=np.int8)
        st, keep = p._struct()
        self._chk(self.lib.orc_update_fields(C.byref(st), _d(f), _b(s), _i64(flip), C.c_int8(old)))
        return f

    def qubo_to_ising(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        n = Q.shape[0]
        h, J, c = np.zeros(n), np.zeros((n, n)), C.c_double()
        self._chk(self.lib.orc_qubo_to_ising(_i64(n), _d(Q), _d(h), _d(J), C.byref(c)))
        return h, J, c.value

    def sa_run(self, p: Problem, betas, sweeps_per_beta: int, seed: int):
        b = np.ascontiguousarray(betas, dtype=np.float64)
        best, be, tr = np.zeros(p.n, np.int8), C.c_double(), np.zeros(b.shape[0])
        st, keep = p._struct()
        self._chk(self.lib.orc_sa_run(C.byref(st), _i64(b.shape[0]), _d(b), _i64(sweeps_per_beta),
                                      C.c_uint64(seed), _b(best), C.byref(be), _d(tr)))
        return dict(best_state=best, best_energy=be.value, energy_trace=tr)

    def replica_run(self, p: Problem, betas, sweeps_per_beta: int, replicas: int, seed: int):
        """ReplicaAnnealer::run (replica_annealer.cpp:46-142) after set_seed(seed)."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        S, R, n = b.shape[0], replicas, p.n
        bs, be = np.zeros((R, n), np.int8), np.zeros(R)
        et, mt = np.zeros((R, S)), np.zeros((R, S))
        gb, gbe = np.zeros(n, np.int8), C.c_double()
        ae, am = np.zeros(S), np.zeros(S)
        st, keep = p._struct()
        self._chk(self.lib.orc_replica_run(C.byref(st), _i64(S), _d(b), _i64(sweeps_per_beta), _i64(R),
                                           C.c_uint64(seed), _b(bs), _d(be), _d(et), _d(mt), _b(gb),
                                           C.byref(gbe), _d(ae), _d(am)))
        return dict(best_states=bs, best_energies=be, energy_traces=et, magnetization_traces=mt,
                    global_best_state=gb, global_best_energy=gbe.value, average_energy_trace=ae,
                    average_magnetization_trace=am)

    def pt_run(self, p: Problem, betas, sweeps_per_step: int, steps: int, swap_interval: int, seed: int):
        """ParallelTemperingAnnealer::run (parallel_tempering.cpp:36-122) after set_seed(seed)."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        R, n = b.shape[0], p.n
        fs, fe = np.zeros((R, n), np.int8), np.zeros(R)
        best, be = np.zeros(n, np.int8), C.c_double()
        at, sw = np.zeros(steps), np.zeros(steps)
        st, keep = p._struct()
        self._chk(self.lib.orc_pt_run(C.byref(st), _i64(R), _d(b), _i64(sweeps_per_step), _i64(steps),
                                      _i64(swap_interval), C.c_uint64(seed), _b(fs), _d(fe), _b(best),
                                      C.byref(be), _d(at), _d(sw)))
        return dict(final_states=fs, final_energies=fe, best_state=best, best_energy=be.value,
                    average_energy_trace=at, swap_acceptance_trace=sw)

    def sa_observe_run(self, p: Problem, betas, sweeps_per_beta: int, seed: int):
        """Annealer::run with an observer that also records every sweep (annealer.cpp:33-75)."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        S, n = b.shape[0], p.n
        best, be, tr, mg = np.zeros(n, np.int8), C.c_double(), np.zeros(S), np.zeros(S)
        se, ss = np.zeros(S * sweeps_per_beta), np.zeros((S * sweeps_per_beta, n), np.int8)
        st, keep = p._struct()
        self._chk(self.lib.orc_sa_observe_run(C.byref(st), _i64(S), _d(b), _i64(sweeps_per_beta), C.c_uint64(seed),
                                              _b(best), C.byref(be), _d(tr), _d(mg), _d(se), _b(ss)))
        return dict(best_state=best, best_energy=be.value, energy_trace=tr, magnetization_trace=mg,
                    sweep_energies=se, sweep_states=ss)

    def sqa_observe_run(self, p: Problem, betas, gammas, trotter_slices: int, replicas: int, sweeps_per_beta: int,
                        worldline_sweeps: int = 0, cluster_sweeps: int = 0, tracing: bool = False, seed: int = 0):
        """SQAAnnealer::run with an SQAObserver; tracing: also an SQASweepObserver (sqa_annealer.cpp:317-453)."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        g = np.ascontiguousarray(gammas, dtype=np.float64)
        S, n, R, P = b.shape[0], p.n, replicas, trotter_slices
        K = S * R * (sweeps_per_beta + worldline_sweeps + cluster_sweeps) if tracing else 0
        best, be, tr, mg = np.zeros(n, np.int8), C.c_double(), np.zeros(S), np.zeros(S)
        meta, av, rep = np.zeros((max(K, 1), 4), np.int64), np.zeros(max(K, 1)), np.zeros((max(K, 1), R))
        states = np.zeros((max(K, 1), R * P * n), np.int8)
        st, keep = p._struct()
        self._chk(self.lib.orc_sqa_observe_run(
            C.byref(st), _i64(S), _d(b), _d(g), _i64(P), _i64(R), _i64(sweeps_per_beta), _i64(worldline_sweeps),
            _i64(cluster_sweeps), C.c_int32(1 if tracing else 0), C.c_uint64(seed), _b(best), C.byref(be), _d(tr),
            _d(mg), meta.ctypes.data_as(C.POINTER(C.c_int64)), _d(av), _d(rep), _b(states)))
        out = dict(best_state=best, best_energy=be.value, energy_trace=tr, magnetization_trace=mg)
        if tracing:
            out.update(sweep_meta=meta[:K], sweep_avg=av[:K], sweep_rep=rep[:K], sweep_states=states[:K])
        return out

    def sqa_opt_run(self, p: Problem, trotter_slices: int, replicas: int, beta: float, j_perp_start: float,
                    j_perp_end: float, eps_tilde: float, alpha: float, num_steps: int, sweeps_per_step: int,
                    worldline_sweeps: int = 0, cluster_sweeps: int = 0, calib_probes: int = 12,
                    calib_sweeps: int = 10, beta_ramp_fraction: float = 0.3, beta_ramp_start: float = 0.1,
                    seed: int = 0):
        """SQAAnnealer::run_optimal (sqa_annealer.cpp:499-952) after set_seed(seed)."""
        n, S = p.n, num_steps
        best, be = np.zeros(n, np.int8), C.c_double()
        tr, jt, ct, sc = np.zeros(S), np.zeros(S), np.zeros(S), np.zeros(4)
        st, keep = p._struct()
        self._chk(self.lib.orc_sqa_opt_run(
            C.byref(st), _i64(trotter_slices), _i64(replicas), C.c_double(beta), C.c_double(j_perp_start),
            C.c_double(j_perp_end), C.c_double(eps_tilde), C.c_double(alpha), _i64(num_steps),
            _i64(sweeps_per_step), _i64(worldline_sweeps), _i64(cluster_sweeps), _i64(calib_probes),
            _i64(calib_sweeps), C.c_double(beta_ramp_fraction), C.c_double(beta_ramp_start), C.c_uint64(seed),
            _b(best), C.byref(be), _d(tr), _d(jt), _d(ct), _d(sc)))
        return dict(best_state=best, best_energy=be.value, energy_trace=tr, j_perp_trace=jt, chi_B_trace=ct,
                    calibrated_eps_tilde=sc[0], j_perp_start=sc[1], resolved_j_perp_end=sc[2], final_j_perp=sc[3])

    def sqapt_opt_run(self, p: Problem, betas, gammas, trotter_slices: int, num_steps: int, sweeps_per_step: int,
                      worldline_sweeps: int = 0, eps_tilde: float = 0.0, alpha: float = 15.0 / 14.0,
                      j_perp_end: float = 0.0, cluster_sweeps: int = 0, swap_interval: int = 1,
                      continuous_time_slices: int = 0, calib_probes: int = 12, calib_sweeps: int = 10, seed: int = 0):
        """SQAParallelTemperingAnnealer::run_optimal (sqa_parallel_tempering.cpp:348-882) after set_seed(seed)."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        g = np.ascontiguousarray(gammas, dtype=np.float64)
        R, n, S = b.shape[0], p.n, num_steps
        fs, fe = np.zeros((R, n), np.int8), np.zeros(R)
        best, be = np.zeros(n, np.int8), C.c_double()
        at, sw, jt, ct, sc = np.zeros(S), np.zeros(S), np.zeros(S), np.zeros(S), np.zeros(4)
        st, keep = p._struct()
        self._chk(self.lib.orc_sqapt_opt_run(
            C.byref(st), _i64(R), _d(b), _d(g), _i64(trotter_slices), _i64(num_steps), _i64(sweeps_per_step),
            _i64(worldline_sweeps), C.c_double(eps_tilde), C.c_double(alpha), C.c_double(j_perp_end),
            _i64(cluster_sweeps), _i64(swap_interval), _i64(continuous_time_slices), _i64(calib_probes),
            _i64(calib_sweeps), C.c_uint64(seed), _b(fs), _d(fe), _b(best), C.byref(be), _d(at), _d(sw), _d(jt),
            _d(ct), _d(sc)))
        return dict(final_states=fs, final_energies=fe, best_state=best, best_energy=be.value,
                    average_energy_trace=at, swap_acceptance_trace=sw, j_perp_trace=jt, chi_B_trace=ct,
                    calibrated_eps_tilde=sc[0], j_perp_start=sc[1], resolved_j_perp_end=sc[2], final_j_perp=sc[3])

    def sqa_run(self, p: Problem, betas, gammas, trotter_slices: int, replicas: int,
                sweeps_per_beta: int, worldline_sweeps: int, cluster_sweeps: int = 0,
                continuous_time_slices: int = 0, seed: int = 0):
        b = np.ascontiguousarray(betas, dtype=np.float64)
        g = np.ascontiguousarray(gammas, dtype=np.float64)
        best, be, tr = np.zeros(p.n, np.int8), C.c_double(), np.zeros(b.shape[0])
        st, keep = p._struct()
        self._chk(self.lib.orc_sqa_run(
            C.byref(st), _i64(b.shape[0]), _d(b), _d(g), _i64(trotter_slices), _i64(replicas),
            _i64(sweeps_per_beta), _i64(worldline_sweeps), _i64(cluster_sweeps),
            _i64(continuous_time_slices), C.c_uint64(seed), _b(best), C.byref(be), _d(tr)))
        return dict(best_state=best, best_energy=be.value, energy_trace=tr)

    def sqa_run_repeat(self, p: Problem, runs: int, betas, gammas, trotter_slices: int, replicas: int,
                       sweeps_per_beta: int, worldline_sweeps: int, cluster_sweeps: int = 0,
                       continuous_time_slices: int = 0, seed: int = 0):
        """set_seed(seed), then `runs` consecutive run() calls on one annealer; the last run's
        outputs."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        g = np.ascontiguousarray(gammas, dtype=np.float64)
        best, be, tr = np.zeros(p.n, np.int8), C.c_double(), np.zeros(b.shape[0])
        st, keep = p._struct()
        self._chk(self.lib.orc_sqa_run_repeat(
            C.byref(st), _i64(runs), _i64(b.shape[0]), _d(b), _d(g), _i64(trotter_slices), _i64(replicas),
            _i64(sweeps_per_beta), _i64(worldline_sweeps), _i64(cluster_sweeps),
            _i64(continuous_time_slices), C.c_uint64(seed), _b(best), C.byref(be), _d(tr)))
        return dict(best_state=best, best_energy=be.value, energy_trace=tr)

    def sqapt_run(self, p: Problem, betas, gammas, trotter_slices: int, sweeps_per_step: int,
                  worldline_sweeps: int, steps: int, swap_interval: int = 1, cluster_sweeps: int = 0,
                  continuous_time_slices: int = 0, seed: int = 0):
        b = np.ascontiguousarray(betas, dtype=np.float64)
        g = np.ascontiguousarray(gammas, dtype=np.float64)
        R = b.shape[0]
        fs, fe = np.zeros((R, p.n), np.int8), np.zeros(R)
        best, be = np.zeros(p.n, np.int8), C.c_double()
        at, sw = np.zeros(steps), np.zeros(steps)
        st, keep = p._struct()
        self._chk(self.lib.orc_sqapt_run(
            C.byref(st), _i64(R), _d(b), _d(g), _i64(trotter_slices), _i64(sweeps_per_step),
            _i64(worldline_sweeps), _i64(steps), _i64(swap_interval), _i64(cluster_sweeps),
            _i64(continuous_time_slices), C.c_uint64(seed), _b(fs), _d(fe), _b(best), C.byref(be),
            _d(at), _d(sw)))
        return dict(final_states=fs, final_energies=fe, best_state=best, best_energy=be.value,
                    average_energy_trace=at, swap_acceptance_trace=sw)

    def sqa_chi_run(self, p: Problem, trotter_slices: int, replicas: int, beta: float,
                    gamma_start: float, gamma_end: float, num_steps: int = 100,
                    sweeps_per_step: int = 20, scan_points: int = 16, scan_sweeps: int = 30,
                    scan_burn: int = 10, chi_floor_fraction: float = 1e-6, driver_A0: float = 0.0,
                    beta_ramp_fraction: float = 0.3, beta_ramp_start: float = 0.1, seed: int = 0):
        best, be = np.zeros(p.n, np.int8), C.c_double()
        tr, bsch, gsch, ssch = (np.zeros(num_steps) for _ in range(4))
        ss, sg, sc = (np.zeros(scan_points) for _ in range(3))
        sc5 = np.zeros(5)
        st, keep = p._struct()
        self._chk(self.lib.orc_sqa_chi_run(
            C.byref(st), _i64(trotter_slices), _i64(replicas), C.c_double(beta),
            C.c_double(gamma_start), C.c_double(gamma_end), _i64(num_steps), _i64(sweeps_per_step),
            _i64(scan_points), _i64(scan_sweeps), _i64(scan_burn), C.c_double(chi_floor_fraction),
            C.c_double(driver_A0), C.c_double(beta_ramp_fraction), C.c_double(beta_ramp_start),
            C.c_uint64(seed), _b(best), C.byref(be), _d(tr), _d(bsch), _d(gsch), _d(ssch), _d(ss),
            _d(sg), _d(sc), _d(sc5)))
        return dict(best_state=best, best_energy=be.value, energy_trace=tr, beta_schedule=bsch,
                    gamma_schedule=gsch, s_schedule=ssch, scan_s=ss, scan_gamma=sg, scan_chi_B=sc,
                    s_star=sc5[0], gamma_star=sc5[1], j_perp_star=sc5[2], chi_floor=sc5[3],
                    driver_A0=sc5[4])
