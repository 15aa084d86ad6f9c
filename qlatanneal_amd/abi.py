"""ctypes binding of the C-ABI in include/qanneal_hip.h (libqanneal_hip.so).

This is the binding a Python caller of the reference would add to reach the device path without
the C++ classes (INTEGRATION.md shows it beside the pybind11 route). It fails loudly when the
library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libqanneal_hip.so")

QA_OK, QA_EINVAL, QA_ENOMEM, QA_EDEVICE, QA_EUNSUPPORTED = range(5)
MODE_FP32, MODE_FP64_VALIDATION = 0, 1

_i64, _u64, _i32, _dbl = C.c_int64, C.c_uint64, C.c_int32, C.c_double
_dp, _bp, _lp, _up, _u32p = (C.POINTER(C.c_double), C.POINTER(C.c_int8), C.POINTER(C.c_int64),
                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint32))


class SaParams(C.Structure):
    _fields_ = [("steps", _i64), ("betas", _dp), ("sweeps_per_beta", _i64), ("reads", _i64),
                ("mode", _i32), ("seed", _u64), ("init_spins", _bp), ("uniforms", _dp),
                ("uniforms_per_read", _i64), ("read_betas", _dp), ("init_states", _bp), ("init_energies", _dp),
                ("counter_base", _u64)]


class SaResult(C.Structure):
    _fields_ = [("best_states", _bp), ("best_energies", _dp), ("energy_trace", _dp),
                ("uniforms_used", _lp), ("kernel_ms", _dp), ("final_states", _bp), ("final_energies", _dp),
                ("magnetization_trace", _dp), ("best_energy_trace", _dp)]


class SqaParams(C.Structure):
    _fields_ = [("steps", _i64), ("betas", _dp), ("gammas", _dp), ("trotter_slices", _i64),
                ("replicas", _i64), ("sweeps_per_beta", _i64), ("worldline_sweeps", _i64),
                ("cluster_sweeps", _i64), ("mode", _i32), ("seed", _u64), ("init_spins", _bp),
                ("replica_seeds", _up), ("refresh_interval", _i64), ("on_step", C.c_void_p),
                ("on_sweep", C.c_void_p), ("callback_user", C.c_void_p), ("trace_engine", C.c_char_p)]


class SqaResult(C.Structure):
    _fields_ = [("best_state", _bp), ("best_energy", _dp), ("energy_trace", _dp),
                ("kernel_ms", _dp), ("accept_rate", _dp), ("field_drift", _dp),
                ("trace_engine_out", C.c_char_p), ("trace_engine_out_len", _i64)]


class SqaptParams(C.Structure):
    _fields_ = [("rungs", _i64), ("betas", _dp), ("gammas", _dp), ("ladders", _i64),
                ("trotter_slices", _i64), ("sweeps_per_step", _i64), ("worldline_sweeps", _i64),
                ("cluster_sweeps", _i64), ("steps", _i64), ("swap_interval", _i64), ("mode", _i32),
                ("seed", _u64), ("init_spins", _bp), ("rung_seeds", _up), ("swap_uniforms", _dp),
                ("swap_uniforms_len", _i64), ("refresh_interval", _i64), ("step_betas", _dp),
                ("step_gammas", _dp)]


class SqaptResult(C.Structure):
    _fields_ = [("final_states", _bp), ("final_energies", _dp), ("best_state", _bp),
                ("best_energy", _dp), ("average_energy_trace", _dp), ("swap_acceptance_trace", _dp),
                ("swap_uniforms_used", _lp), ("kernel_ms", _dp), ("field_drift", _dp), ("sweep_ms", _dp),
                ("accept_rate", _dp)]


class ChiParams(C.Structure):
    _fields_ = [("trotter_slices", _i64), ("replicas", _i64), ("beta", _dbl), ("gamma_start", _dbl),
                ("gamma_end", _dbl), ("num_steps", _i64), ("sweeps_per_step", _i64),
                ("scan_points", _i64), ("scan_sweeps", _i64), ("scan_burn", _i64),
                ("chi_floor_fraction", _dbl), ("driver_A0", _dbl), ("beta_ramp_fraction", _dbl),
                ("beta_ramp_start", _dbl), ("mode", _i32), ("seed", _u64), ("init_spins", _bp),
                ("cell_seeds", _up), ("chi_override", _dp), ("surrogate", C.c_void_p),
                ("surrogate_max_spread", _dbl)]


class ChiResult(C.Structure):
    _fields_ = [("best_state", _bp), ("best_energy", _dp), ("energy_trace", _dp),
                ("beta_schedule", _dp), ("gamma_schedule", _dp), ("s_schedule", _dp),
                ("scan_s", _dp), ("scan_gamma", _dp), ("scan_chi_B", _dp), ("scalars", _dp),
                ("kernel_ms", _dp), ("surrogate_used", C.POINTER(_i32)), ("surrogate_spread", _dp)]


class MlpTrainParams(C.Structure):
    _fields_ = [("epochs", _i64), ("batch", _i64), ("lr", _dbl), ("beta1", _dbl), ("beta2", _dbl), ("eps", _dbl),
                ("weight_decay", _dbl), ("seed", _u64)]


class BatchParams(C.Structure):
    _fields_ = [("steps", _i64), ("betas", _dp), ("gammas", _dp), ("trotter_slices", _i64),
                ("sweeps_per_beta", _i64), ("worldline_sweeps", _i64), ("cluster_sweeps", _i64),
                ("reads", _lp), ("seeds", _up), ("autotune", _i32), ("use_graph", _i32)]


class BatchResult(C.Structure):
    _fields_ = [("best_states", _bp), ("best_energies", _dp), ("kernel_ms", _dp), ("attempts", _lp)]


class QaError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


class QaInvalidArgument(QaError, ValueError):
    pass


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build()); "
                              "qlatanneal_amd has no CPU fallback for the annealing sweep")
        L = C.CDLL(LIB_PATH)
        L.qa_last_error.restype = C.c_char_p
        for name in ("qa_ctx_create", "qa_ctx_destroy", "qa_problem_dense", "qa_problem_sparse",
                     "qa_problem_hubo", "qa_problem_destroy", "qa_energy", "qa_local_fields", "qa_sa_run", "qa_sqa_run",
                     "qa_sqa_run_noalloc",
                     "qa_sqapt_run", "qa_sqa_chi_run", "qa_chi_schedule", "qa_sqa_batch_run", "qa_chi_pilot",
                     "qa_chi_features",
                     "qa_mlp_create", "qa_mlp_destroy", "qa_mlp_info", "qa_mlp_get_state", "qa_mlp_set_state",
                     "qa_mlp_forward", "qa_mlp_loss_grad", "qa_mlp_train"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != QA_OK:
        msg = lib().qa_last_error().decode()
        if status == QA_EINVAL:
            raise QaInvalidArgument(status, msg)
        raise QaError(status, msg)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _b(a):
    return a.ctypes.data_as(_bp) if a is not None else None


class Context:
    """One GPU (qa_ctx)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        self.device = device
        check(lib().qa_ctx_create(C.c_int(device), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().qa_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceProblem:
    """A problem model uploaded once (qa_problem)."""

    def __init__(self, ctx: Context, handle: C.c_void_p, n: int):
        self.ctx, self.h, self.n = ctx, handle, n

    @staticmethod
    def dense(ctx: Context, h, J, c: float = 0.0) -> "DeviceProblem":
        h = np.ascontiguousarray(h, np.float64)
        J = np.ascontiguousarray(J, np.float64)
        n = h.shape[0]
        if J.shape != (n, n):
            raise ValueError("J must be n x n")
        out = C.c_void_p()
        check(lib().qa_problem_dense(ctx.h, _i64(n), _d(h), _d(J), _dbl(c), C.byref(out)))
        return DeviceProblem(ctx, out, n)

    @staticmethod
    def sparse(ctx: Context, h, edges, c: float = 0.0) -> "DeviceProblem":
        h = np.ascontiguousarray(h, np.float64)
        n = h.shape[0]
        ei = np.array([int(e[0]) for e in edges], np.int64)
        ej = np.array([int(e[1]) for e in edges], np.int64)
        ev = np.array([float(e[2]) for e in edges], np.float64)
        out = C.c_void_p()
        check(lib().qa_problem_sparse(ctx.h, _i64(n), _d(h), _i64(len(edges)), ei.ctypes.data_as(_lp),
                                      ej.ctypes.data_as(_lp), _d(ev), _dbl(c), C.byref(out)))
        return DeviceProblem(ctx, out, n)

    @staticmethod
    def hubo(ctx: Context, n: int, h, term_offset, term_vars, term_coeff, c: float = 0.0) -> "DeviceProblem":
        h = np.ascontiguousarray(h, np.float64)
        off = np.ascontiguousarray(term_offset, np.int64)
        tv = np.ascontiguousarray(term_vars if len(term_vars) else [0], np.uint32)
        tc = np.ascontiguousarray(term_coeff if len(term_coeff) else [0.0], np.float64)
        out = C.c_void_p()
        check(lib().qa_problem_hubo(ctx.h, _i64(n), _d(h), _i64(len(off) - 1), off.ctypes.data_as(_lp),
                                    tv.ctypes.data_as(_u32p), _d(tc), _dbl(c), C.byref(out)))
        return DeviceProblem(ctx, out, n)

    def energy(self, spins) -> np.ndarray:
        s = np.ascontiguousarray(spins, np.int8).reshape(-1, self.n)
        out = np.zeros(s.shape[0])
        check(lib().qa_energy(self.h, _i64(s.shape[0]), _b(s), _d(out)))
        return out

    def local_fields(self, spins) -> np.ndarray:
        """FP32 local fields of each state (matrix cores for symmetric dense couplings)."""
        s = np.ascontiguousarray(spins, np.int8).reshape(-1, self.n)
        out = np.zeros(s.shape, np.float32)
        check(lib().qa_local_fields(self.h, _i64(s.shape[0]), _b(s),
                                    out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def close(self):
        if self.h:
            lib().qa_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sa_run(prob: DeviceProblem, betas, sweeps_per_beta: int, reads: int = 1, mode: int = MODE_FP32,
           seed: int = 0, init_spins=None, uniforms=None, read_betas=None, init_states=None, init_energies=None,
           counter_base: int = 0):
    betas = np.ascontiguousarray(betas, np.float64)
    steps, n = betas.shape[0], prob.n
    p = SaParams(steps, _d(betas), sweeps_per_beta, reads, mode, seed & (2**64 - 1))
    p.counter_base = counter_base & (2**64 - 1)
    keep = []
    for name, arr, dt in (("read_betas", read_betas, np.float64), ("init_states", init_states, np.int8),
                          ("init_energies", init_energies, np.float64)):
        if arr is not None:
            arr = np.ascontiguousarray(arr, dt)
            keep.append(arr)
            setattr(p, name, _b(arr) if dt == np.int8 else _d(arr))
    if init_spins is not None:
        init_spins = np.ascontiguousarray(init_spins, np.int8)
        keep.append(init_spins)
        p.init_spins = _b(init_spins)
    if uniforms is not None:
        uniforms = np.ascontiguousarray(uniforms, np.float64).reshape(reads, -1)
        keep.append(uniforms)
        p.uniforms = _d(uniforms)
        p.uniforms_per_read = uniforms.shape[1]
    bs, be, tr = np.zeros((reads, n), np.int8), np.zeros(reads), np.zeros((reads, steps))
    used, ms = np.zeros(reads, np.int64), np.zeros(1)
    fs, fe, mg = np.zeros((reads, n), np.int8), np.zeros(reads), np.zeros((reads, steps))
    bt = np.zeros((reads, steps + 1))
    r = SaResult(_b(bs), _d(be), _d(tr), used.ctypes.data_as(_lp), _d(ms), _b(fs), _d(fe), _d(mg), _d(bt))
    check(lib().qa_sa_run(prob.ctx.h, prob.h, C.byref(p), C.byref(r)))
    return dict(best_states=bs, best_energies=be, energy_traces=tr, uniforms_used=used, kernel_ms=ms[0],
                final_states=fs, final_energies=fe, magnetization_traces=mg, best_energy_traces=bt)


def sqa_run(prob: DeviceProblem, betas, gammas, trotter_slices: int, replicas: int, sweeps_per_beta: int,
            worldline_sweeps: int = 0, cluster_sweeps: int = 0, mode: int = MODE_FP32, seed: int = 0,
            init_spins=None, replica_seeds=None, refresh_interval: int = 0):
    betas = np.ascontiguousarray(betas, np.float64)
    gammas = np.ascontiguousarray(gammas, np.float64)
    steps, n = betas.shape[0], prob.n
    p = SqaParams(steps, _d(betas), _d(gammas), trotter_slices, replicas, sweeps_per_beta,
                  worldline_sweeps, cluster_sweeps, mode, seed & (2**64 - 1))
    p.refresh_interval = refresh_interval
    keep = []
    if init_spins is not None:
        init_spins = np.ascontiguousarray(init_spins, np.int8)
        keep.append(init_spins)
        p.init_spins = _b(init_spins)
    if replica_seeds is not None:
        replica_seeds = np.ascontiguousarray(replica_seeds, np.uint64)
        keep.append(replica_seeds)
        p.replica_seeds = replica_seeds.ctypes.data_as(_up)
    bs, be, tr, ms, ar = np.zeros(n, np.int8), np.zeros(1), np.zeros(steps), np.zeros(1), np.zeros(1)
    dr = np.zeros(2)
    r = SqaResult(_b(bs), _d(be), _d(tr), _d(ms), _d(ar), _d(dr))
    check(lib().qa_sqa_run(prob.ctx.h, prob.h, C.byref(p), C.byref(r)))
    return dict(best_state=bs, best_energy=float(be[0]), energy_trace=tr, kernel_ms=float(ms[0]),
                accept_rate=float(ar[0]), field_drift=dr)


def sqa_run_noalloc(prob: DeviceProblem, betas, gammas, trotter_slices: int, replicas: int, sweeps_per_beta: int,
                    worldline_sweeps: int = 0, cluster_sweeps: int = 0, seed: int = 0):
    """FP32 SQA run through qa_sqa_run_noalloc: no device allocation (a static per-device scratch
    array of 64 MiB holds the run's buffers); results equal sqa_run(..., mode=MODE_FP32)."""
    betas = np.ascontiguousarray(betas, np.float64)
    gammas = np.ascontiguousarray(gammas, np.float64)
    steps, n = betas.shape[0], prob.n
    p = SqaParams(steps, _d(betas), _d(gammas), trotter_slices, replicas, sweeps_per_beta,
                  worldline_sweeps, cluster_sweeps, MODE_FP32, seed & (2**64 - 1))
    bs, be, tr, ms, ar = np.zeros(n, np.int8), np.zeros(1), np.zeros(steps), np.zeros(1), np.zeros(1)
    dr = np.zeros(2)
    r = SqaResult(_b(bs), _d(be), _d(tr), _d(ms), _d(ar), _d(dr))
    check(lib().qa_sqa_run_noalloc(prob.ctx.h, prob.h, C.byref(p), C.byref(r)))
    return dict(best_state=bs, best_energy=float(be[0]), energy_trace=tr, kernel_ms=float(ms[0]),
                accept_rate=float(ar[0]))


def sqapt_run(prob: DeviceProblem, betas, gammas, trotter_slices: int, sweeps_per_step: int,
              worldline_sweeps: int, steps: int, swap_interval: int = 1, cluster_sweeps: int = 0,
              ladders: int = 1, mode: int = MODE_FP32, seed: int = 0, init_spins=None, rung_seeds=None,
              swap_uniforms=None, refresh_interval: int = 0, step_betas=None, step_gammas=None):
    """SQA parallel tempering (qa_sqapt_run). step_betas / step_gammas ([steps, rungs], FP32 mode)
    give every PT step its own ladder; None keeps betas/gammas fixed."""
    betas = np.ascontiguousarray(betas, np.float64)
    gammas = np.ascontiguousarray(gammas, np.float64)
    R, n = betas.shape[0], prob.n
    p = SqaptParams(R, _d(betas), _d(gammas), ladders, trotter_slices, sweeps_per_step, worldline_sweeps,
                    cluster_sweeps, steps, swap_interval, mode, seed & (2**64 - 1))
    p.refresh_interval = refresh_interval
    keep = []
    if (step_betas is None) != (step_gammas is None):
        raise ValueError("step_betas and step_gammas go together")
    if step_betas is not None:
        step_betas = np.ascontiguousarray(step_betas, np.float64)
        step_gammas = np.ascontiguousarray(step_gammas, np.float64)
        if step_betas.shape != (steps, R) or step_gammas.shape != (steps, R):
            raise ValueError(f"step_betas / step_gammas must have shape ({steps}, {R})")
        keep += [step_betas, step_gammas]
        p.step_betas = _d(step_betas)
        p.step_gammas = _d(step_gammas)
    if init_spins is not None:
        init_spins = np.ascontiguousarray(init_spins, np.int8)
        keep.append(init_spins)
        p.init_spins = _b(init_spins)
    if rung_seeds is not None:
        rung_seeds = np.ascontiguousarray(rung_seeds, np.uint64)
        keep.append(rung_seeds)
        p.rung_seeds = rung_seeds.ctypes.data_as(_up)
    if swap_uniforms is not None:
        swap_uniforms = np.ascontiguousarray(swap_uniforms, np.float64)
        keep.append(swap_uniforms)
        p.swap_uniforms = _d(swap_uniforms)
        p.swap_uniforms_len = swap_uniforms.shape[0]
    L = ladders
    fs, fe = np.zeros((L, R, n), np.int8), np.zeros((L, R))
    bs, be = np.zeros((L, n), np.int8), np.zeros(L)
    at, st = np.zeros((L, steps)), np.zeros((L, steps))
    used, ms, dr, sw, ar = np.zeros(1, np.int64), np.zeros(1), np.zeros(2), np.zeros(1), np.full(1, np.nan)
    r = SqaptResult(_b(fs), _d(fe), _b(bs), _d(be), _d(at), _d(st), used.ctypes.data_as(_lp), _d(ms), _d(dr),
                    _d(sw) if mode == MODE_FP32 else None, _d(ar) if mode == MODE_FP32 else None)
    check(lib().qa_sqapt_run(prob.ctx.h, prob.h, C.byref(p), C.byref(r)))
    return dict(final_states=fs, final_energies=fe, best_state=bs, best_energy=be, average_energy_trace=at,
                swap_acceptance_trace=st, swap_uniforms_used=int(used[0]), kernel_ms=float(ms[0]),
                field_drift=dr, sweep_ms=float(sw[0]), accept_rate=float(ar[0]))


def sqa_chi_run(prob: DeviceProblem, trotter_slices: int, replicas: int, beta: float, gamma_start: float,
                gamma_end: float, num_steps: int = 100, sweeps_per_step: int = 20, scan_points: int = 16,
                scan_sweeps: int = 30, scan_burn: int = 10, chi_floor_fraction: float = 1e-6,
                driver_A0: float = 0.0, beta_ramp_fraction: float = 0.3, beta_ramp_start: float = 0.1,
                mode: int = MODE_FP32, seed: int = 0, init_spins=None, cell_seeds=None, chi_override=None,
                surrogate: "Optional[Mlp]" = None, surrogate_max_spread: float = 0.05):
    p = ChiParams(trotter_slices, replicas, beta, gamma_start, gamma_end, num_steps, sweeps_per_step, scan_points,
                  scan_sweeps, scan_burn, chi_floor_fraction, driver_A0, beta_ramp_fraction, beta_ramp_start,
                  mode, seed & (2**64 - 1))
    keep = []
    if init_spins is not None:
        init_spins = np.ascontiguousarray(init_spins, np.int8)
        keep.append(init_spins)
        p.init_spins = _b(init_spins)
    if cell_seeds is not None:
        cell_seeds = np.ascontiguousarray(cell_seeds, np.uint64)
        keep.append(cell_seeds)
        p.cell_seeds = cell_seeds.ctypes.data_as(_up)
    if chi_override is not None:
        chi_override = np.ascontiguousarray(chi_override, np.float64)
        keep.append(chi_override)
        p.chi_override = _d(chi_override)
    if surrogate is not None:
        p.surrogate = surrogate.h
        p.surrogate_max_spread = float(surrogate_max_spread)
    n = prob.n
    bs, be = np.zeros(n, np.int8), np.zeros(1)
    tr, bsch, gsch, ssch = (np.zeros(num_steps) for _ in range(4))
    ss, sg, sc = (np.zeros(scan_points) for _ in range(3))
    sc5, ms = np.zeros(5), np.zeros(1)
    used, spread = np.zeros(1, np.int32), np.zeros(1)
    r = ChiResult(_b(bs), _d(be), _d(tr), _d(bsch), _d(gsch), _d(ssch), _d(ss), _d(sg), _d(sc), _d(sc5), _d(ms),
                  used.ctypes.data_as(C.POINTER(_i32)), _d(spread))
    check(lib().qa_sqa_chi_run(prob.ctx.h, prob.h, C.byref(p), C.byref(r)))
    return dict(best_state=bs, best_energy=float(be[0]), energy_trace=tr, beta_schedule=bsch, gamma_schedule=gsch,
                s_schedule=ssch, scan_s=ss, scan_gamma=sg, scan_chi_B=sc, s_star=sc5[0], gamma_star=sc5[1],
                j_perp_star=sc5[2], chi_floor=sc5[3], driver_A0=sc5[4], kernel_ms=float(ms[0]),
                surrogate_used=bool(used[0]), surrogate_spread=float(spread[0]))


def chi_schedule(prob: DeviceProblem, trotter_slices: int, replicas: int, beta: float, gamma_start: float,
                 gamma_end: float, num_steps: int = 100, scan_points: int = 16, scan_sweeps: int = 30,
                 scan_burn: int = 10, chi_floor_fraction: float = 1e-6, driver_A0: float = 0.0,
                 beta_ramp_fraction: float = 0.3, beta_ramp_start: float = 0.1, seed: int = 0, chi_override=None,
                 surrogate: "Optional[Mlp]" = None, surrogate_max_spread: float = 0.05):
    """run_chi's chi_B profile and schedules without its production anneal (qa_chi_schedule, FP32):
    the beta/gamma/s schedules sqa_chi_run would anneal on for the same arguments."""
    p = ChiParams(trotter_slices, replicas, beta, gamma_start, gamma_end, num_steps, 1, scan_points, scan_sweeps,
                  scan_burn, chi_floor_fraction, driver_A0, beta_ramp_fraction, beta_ramp_start, MODE_FP32,
                  seed & (2**64 - 1))
    keep = []
    if chi_override is not None:
        chi_override = np.ascontiguousarray(chi_override, np.float64)
        keep.append(chi_override)
        p.chi_override = _d(chi_override)
    if surrogate is not None:
        p.surrogate = surrogate.h
        p.surrogate_max_spread = float(surrogate_max_spread)
    bsch, gsch, ssch = (np.zeros(num_steps) for _ in range(3))
    ss, sg, sc = (np.zeros(scan_points) for _ in range(3))
    sc5, ms = np.zeros(5), np.zeros(1)
    used, spread = np.zeros(1, np.int32), np.zeros(1)
    r = ChiResult(None, None, None, _d(bsch), _d(gsch), _d(ssch), _d(ss), _d(sg), _d(sc), _d(sc5), _d(ms),
                  used.ctypes.data_as(C.POINTER(_i32)), _d(spread))
    check(lib().qa_chi_schedule(prob.ctx.h, prob.h, C.byref(p), C.byref(r)))
    return dict(beta_schedule=bsch, gamma_schedule=gsch, s_schedule=ssch, scan_s=ss, scan_gamma=sg, scan_chi_B=sc,
                s_star=sc5[0], gamma_star=sc5[1], j_perp_star=sc5[2], chi_floor=sc5[3], driver_A0=sc5[4],
                kernel_ms=float(ms[0]), surrogate_used=bool(used[0]), surrogate_spread=float(spread[0]))


def chi_pilot(prob: DeviceProblem, trotter_slices: int, replicas: int, beta: float, gamma_start: float,
              gamma_end: float, scan_points: int = 16, scan_sweeps: int = 30, scan_burn: int = 10,
              driver_A0: float = 0.0, seed: int = 0):
    """The pilot chi_B scan of run_chi alone (FP32): scan_s, scan_gamma, scan_chi_B, kernel_ms."""
    p = ChiParams(trotter_slices, replicas, beta, gamma_start, gamma_end, 2, 1, scan_points, scan_sweeps, scan_burn,
                  1e-6, driver_A0, 0.0, 0.0, MODE_FP32, seed & (2**64 - 1))
    ss, sg, sc, ms = np.zeros(scan_points), np.zeros(scan_points), np.zeros(scan_points), np.zeros(1)
    check(lib().qa_chi_pilot(prob.ctx.h, prob.h, C.byref(p), _d(ss), _d(sg), _d(sc), _d(ms)))
    return dict(scan_s=ss, scan_gamma=sg, scan_chi_B=sc, kernel_ms=float(ms[0]))


CHI_NFEATURES = 12


def chi_features(prob: DeviceProblem, beta: float, trotter_slices: int, driver_A0: float, gamma_start: float,
                 gamma_end: float) -> np.ndarray:
    f = np.zeros(CHI_NFEATURES)
    check(lib().qa_chi_features(prob.h, _dbl(beta), _i64(trotter_slices), _dbl(driver_A0), _dbl(gamma_start),
                                _dbl(gamma_end), _d(f)))
    return f


class Mlp:
    """Device MLP ensemble of the chi_B surrogate (qa_mlp_*): members x (n_in -> width tanh x depth
    -> n_out sigmoid)."""

    def __init__(self, ctx: Context, n_in: int, width: int, depth: int, n_out: int, members: int = 1,
                 seed: int = 0):
        self.ctx = ctx
        self.n_in, self.width, self.depth, self.n_out, self.members = n_in, width, depth, n_out, members
        self.h = C.c_void_p()
        check(lib().qa_mlp_create(ctx.h, C.c_int32(n_in), C.c_int32(width), C.c_int32(depth), C.c_int32(n_out),
                                  C.c_int32(members), C.c_uint64(seed & (2**64 - 1)), C.byref(self.h)))
        n, t = C.c_int64(), C.c_int32()
        check(lib().qa_mlp_info(self.h, C.byref(n), C.byref(t)))
        self.params_per_member = int(n.value)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.qa_mlp_destroy(self.h)
            self.h = C.c_void_p()

    @property
    def trained(self) -> bool:
        n, t = C.c_int64(), C.c_int32()
        check(lib().qa_mlp_info(self.h, C.byref(n), C.byref(t)))
        return bool(t.value)

    def get_state(self):
        params = np.zeros(self.members * self.params_per_member, np.float32)
        norm = np.zeros(4 * self.n_in)
        check(lib().qa_mlp_get_state(self.h, params.ctypes.data_as(C.POINTER(C.c_float)), _d(norm)))
        return params.reshape(self.members, self.params_per_member), norm.reshape(4, self.n_in)

    def set_state(self, params=None, norm=None):
        pp = C.POINTER(C.c_float)()
        nn = C.POINTER(C.c_double)()
        if params is not None:
            params = np.ascontiguousarray(params, np.float32).reshape(-1)
            assert params.size == self.members * self.params_per_member
            pp = params.ctypes.data_as(C.POINTER(C.c_float))
        if norm is not None:
            norm = np.ascontiguousarray(norm, np.float64).reshape(-1)
            assert norm.size == 4 * self.n_in
            nn = _d(norm)
        check(lib().qa_mlp_set_state(self.h, pp, nn))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.n_in)
        y = np.zeros((self.members, x.shape[0], self.n_out), np.float32)
        check(lib().qa_mlp_forward(self.h, _i64(x.shape[0]), _d(x), y.ctypes.data_as(C.POINTER(C.c_float))))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.n_in)
        t = np.ascontiguousarray(t, np.float32).reshape(x.shape[0], self.n_out)
        loss = np.zeros(self.members)
        grad = np.zeros((self.members, self.params_per_member), np.float32)
        check(lib().qa_mlp_loss_grad(self.h, _i64(x.shape[0]), _d(x), t.ctypes.data_as(C.POINTER(C.c_float)),
                                     _d(loss), grad.ctypes.data_as(C.POINTER(C.c_float))))
        return loss, grad

    def train(self, x, t, epochs: int, batch: int, lr: float = 1e-3, weight_decay: float = 0.0, seed: int = 0,
              beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.n_in)
        t = np.ascontiguousarray(t, np.float32).reshape(x.shape[0], self.n_out)
        hist = np.zeros(epochs)
        tp = MlpTrainParams(epochs, batch, lr, beta1, beta2, eps, weight_decay, seed & (2**64 - 1))
        check(lib().qa_mlp_train(self.h, _i64(x.shape[0]), _d(x), t.ctypes.data_as(C.POINTER(C.c_float)),
                                 C.byref(tp), _d(hist)))
        return hist


def sqa_batch_run(problems, betas, gammas, trotter_slices: int, sweeps_per_beta: int, reads, seeds,
                  worldline_sweeps: int = 0, cluster_sweeps: int = 0, autotune: bool = False,
                  use_graph: bool = False):
    """Many problems (DeviceProblem, same context) in one persistent launch; per instance the
    FP32 SQA run of sqa_run(problem, ..., replicas=reads[k], seed=seeds[k])."""
    betas = np.ascontiguousarray(betas, np.float64)
    gammas = np.ascontiguousarray(gammas, np.float64)
    count = len(problems)
    reads = np.ascontiguousarray(reads, np.int64)
    seeds = np.ascontiguousarray([int(x) & (2**64 - 1) for x in seeds], np.uint64)
    assert reads.shape == (count,) and seeds.shape == (count,)
    handles = (C.c_void_p * count)(*[p.h.value for p in problems])
    prm = BatchParams(betas.shape[0], _d(betas), _d(gammas), trotter_slices, sweeps_per_beta, worldline_sweeps,
                      cluster_sweeps, reads.ctypes.data_as(_lp), seeds.ctypes.data_as(_up), int(autotune),
                      int(use_graph))
    ns = [p.n for p in problems]
    bs = np.zeros(sum(ns), np.int8)
    be, ms, att = np.zeros(count), np.zeros(1), np.zeros(1, np.int64)
    res = BatchResult(_b(bs), _d(be), _d(ms), att.ctypes.data_as(_lp))
    check(lib().qa_sqa_batch_run(problems[0].ctx.h, _i64(count), handles, C.byref(prm), C.byref(res)))
    offs = np.concatenate([[0], np.cumsum(ns)])
    states = [bs[offs[k]:offs[k + 1]].copy() for k in range(count)]
    return dict(best_states=states, best_energies=be, kernel_ms=float(ms[0]), attempts=int(att[0]))
