"""ctypes wrapper of the batched engine of the C ABI (hs_engine_* in
include/hotspot_amd.h): many thermal models solved by shared device
launches, admitted against a device-memory budget."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import iface

FP64, MIXED = 0, 1
RK4, TAYLOR, BE = 0, 1, 2
QUEUED, STEADY, TRANSIENT, DONE, FAILED = 0, 1, 2, 3, 4


class Options(C.Structure):
    """hs_options_t"""
    _fields_ = [("precision", C.c_int), ("steady_solver", C.c_int), ("transient", C.c_int),
                ("steady_rtol", C.c_double), ("max_iters", C.c_int), ("taylor_tol", C.c_double),
                ("steady_warm_start", C.c_int)]


class JobInfo(C.Structure):
    """hs_job_info_t"""
    _fields_ = [("state", C.c_int), ("steady_iters", C.c_int), ("steady_refinements", C.c_int),
                ("rk4_calls", C.c_int), ("taylor_sweeps", C.c_double), ("wait_s", C.c_double), ("latency_s", C.c_double),
                ("device_bytes", C.c_double), ("taylor_node_bytes", C.c_double), ("held_intervals", C.c_int),
                ("horner_steps_min", C.c_int), ("horner_steps_plan", C.c_int), ("setup_s", C.c_double),
                ("resident_model", C.c_int), ("steady_solves", C.c_int), ("be_iters", C.c_double),
                ("r_convec", C.c_double)]


class EngineStats(C.Structure):
    """hs_engine_stats_t"""
    _fields_ = [("admit_s", C.c_double), ("sync_s", C.c_double), ("launch_s", C.c_double),
                ("wait_s", C.c_double), ("harvest_s", C.c_double), ("steady_chunks", C.c_long),
                ("transient_chunks", C.c_long), ("resident_hits", C.c_long),
                ("steady_chunk_s", C.c_double), ("transient_chunk_s", C.c_double)]


def _bind(L):
    dp = iface.c_double_p
    L.hs_engine_create.restype = C.c_void_p
    L.hs_engine_create.argtypes = [C.c_int, C.c_int]
    L.hs_engine_destroy.argtypes = [C.c_void_p]
    L.hs_engine_destroy.restype = None
    L.hs_engine_set_options.argtypes = [C.c_void_p, C.POINTER(Options)]
    L.hs_engine_set_options.restype = None
    L.hs_engine_set_tick.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.hs_engine_set_tick.restype = None
    L.hs_engine_set_budget.argtypes = [C.c_void_p, C.c_double]
    L.hs_engine_set_budget.restype = None
    L.hs_engine_submit.argtypes = [C.c_void_p, iface.RCModelP, dp, C.c_int, dp, C.c_int, C.c_double]
    L.hs_engine_submit.restype = C.c_int
    L.hs_engine_submit_package.argtypes = [C.c_void_p, iface.RCModelP, dp, C.POINTER(iface.PackageConfig)]
    L.hs_engine_submit_package.restype = C.c_int
    L.hs_engine_admit.argtypes = [C.c_void_p]
    L.hs_engine_admit.restype = C.c_int
    L.hs_engine_tick.argtypes = [C.c_void_p]
    L.hs_engine_tick.restype = C.c_int
    L.hs_engine_poll.argtypes = [C.c_void_p, C.c_int, dp]
    L.hs_engine_poll.restype = C.c_int
    L.hs_engine_result_seq.argtypes = [C.c_void_p, C.c_int, dp]
    L.hs_engine_result_seq.restype = C.c_int
    L.hs_engine_job_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(JobInfo)]
    L.hs_engine_job_info.restype = C.c_int
    L.hs_engine_release.argtypes = [C.c_void_p, C.c_int]
    L.hs_engine_release.restype = None
    L.hs_engine_active.argtypes = [C.c_void_p]
    L.hs_engine_active.restype = C.c_int
    L.hs_engine_get_stats.argtypes = [C.c_void_p, C.POINTER(EngineStats)]
    L.hs_engine_get_stats.restype = None
    L.hs_last_error.argtypes = [C.c_char_p, C.c_int]
    L.hs_last_error.restype = C.c_int


def _dp(a):
    return a.ctypes.data_as(iface.c_double_p)


class Engine:
    """One engine; jobs are models (cases.Run or RC_model_t pointers) with
    power rows, optional start temperatures and a number of intervals."""

    def __init__(self, lib: iface.HotSpotLib, precision=FP64, transient=RK4):
        self.lib = lib
        self.L = lib.lib
        _bind(self.L)
        self.h = self.L.hs_engine_create(precision, transient)
        self._check()
        if not self.h:
            raise RuntimeError("hs_engine_create failed")
        self.meta = {}

    def _check(self):
        buf = C.create_string_buffer(1024)
        if self.L.hs_last_error(buf, 1024):
            raise RuntimeError(buf.value.decode())

    def set_options(self, **kw):
        o = Options(precision=FP64, steady_solver=0, transient=RK4, steady_rtol=1e-12,
                    max_iters=200000, taylor_tol=1e-7, steady_warm_start=0)
        for k, v in kw.items():
            setattr(o, k, v)
        self.L.hs_engine_set_options(self.h, C.byref(o))
        self._check()

    def set_tick(self, intervals, steady_iters):
        self.L.hs_engine_set_tick(self.h, intervals, steady_iters)
        self._check()

    def set_budget(self, nbytes):
        self.L.hs_engine_set_budget(self.h, float(nbytes))
        self._check()

    def submit(self, model, power, init=None, n_steps=0, dt=0.0):
        """power: one block vector or (rows, n) rows; returns the job id."""
        P = np.ascontiguousarray(np.atleast_2d(np.asarray(power, dtype=np.float64)))
        n = P.shape[1]
        I = None if init is None else np.ascontiguousarray(init, dtype=np.float64)
        job = self.L.hs_engine_submit(self.h, model, _dp(P), P.shape[0],
                                      None if I is None else _dp(I), n_steps, dt)
        self._check()
        if job < 0:
            raise RuntimeError("hs_engine_submit refused the job")
        self.meta[job] = (n, n_steps)
        return job

    def submit_package(self, model, power, package):
        """A steady job under the package model (package: iface.PackageConfig);
        with package.natural_convec, the natural-convection fixed point."""
        P = np.ascontiguousarray(np.asarray(power, dtype=np.float64))
        job = self.L.hs_engine_submit_package(self.h, model, _dp(P), C.byref(package))
        self._check()
        if job < 0:
            raise RuntimeError("hs_engine_submit_package refused the job")
        self.meta[job] = (P.shape[-1], 0)
        return job

    def admit(self):
        """Admits what fits; returns the number of jobs still queued."""
        q = self.L.hs_engine_admit(self.h)
        self._check()
        return q

    def tick(self):
        a = self.L.hs_engine_tick(self.h)
        self._check()
        return a

    def active(self):
        return self.L.hs_engine_active(self.h)

    def poll(self, job):
        n, _ = self.meta[job]
        out = np.zeros(n)
        r = self.L.hs_engine_poll(self.h, job, _dp(out))
        self._check()
        return out if r == 1 else None

    def result_seq(self, job):
        n, steps = self.meta[job]
        out = np.zeros((steps, n))
        r = self.L.hs_engine_result_seq(self.h, job, _dp(out))
        self._check()
        return out if r == 1 else None

    def info(self, job):
        i = JobInfo()
        self.L.hs_engine_job_info(self.h, job, C.byref(i))
        self._check()
        return i

    def stats(self):
        st = EngineStats()
        self.L.hs_engine_get_stats(self.h, C.byref(st))
        self._check()
        return st

    def release(self, job):
        self.L.hs_engine_release(self.h, job)
        self._check()
        self.meta.pop(job, None)

    def run(self, max_ticks=10 ** 7):
        """Ticks until every job has finished."""
        for _ in range(max_ticks):
            if self.tick() == 0:
                return
        raise RuntimeError("engine did not drain")

    def close(self):
        if self.h:
            self.L.hs_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
