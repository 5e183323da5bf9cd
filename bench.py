#!/usr/bin/env python3
"""Benchmark of the MI355X grid thermal solver on its headline workload.

Workload (one "step" = one whole job): a 3D-stacked chip on a 256 x 256 grid,
4 silicon dies (0.1 mm, k = 100 W/mK) interleaved with 3 bond layers (10 um,
k = 1.5 W/mK) over the default spreader and sink (9 layers, 589 824 cells),
64 seeded slicing blocks per die with power density lognormal(ln 0.5 W/mm^2,
0.6); a steady solve, then 10 000 transient intervals of 3.333 us with
+-20 % seeded per-interval power jitter, starting from the steady state.

metric: delivered cell-updates/s = cells x (1 steady + intervals) / job time,
for the mixed-precision form (FP32 stencil sweeps inside FP64 refinement /
FP64 residuals, Taylor transient) with inputs resident on the device.  The
plain FP64 GPU form (FP64 Jacobi-CG steady, the reference's adaptive RK4
transient) runs the same job once; the JSON line carries its rate, the
speedup and the largest block-temperature difference between the two forms.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from hotspot_amd import cases, iface  # noqa: E402

PROD_LIB = os.path.join(ROOT, "hotspot_amd", "libhotspot_amd.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libhotspot_ref.so")
HBM_PEAK_GBS = 8000.0
DT = 3.333e-6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("steady_solver", C.c_int), ("transient", C.c_int),
                ("steady_rtol", C.c_double), ("max_iters", C.c_int), ("taylor_tol", C.c_double),
                ("steady_warm_start", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("steady_iters", C.c_int), ("steady_refinements", C.c_int),
                ("steady_rel_resid", C.c_double), ("rk4_calls", C.c_int),
                ("taylor_terms", C.c_int), ("taylor_substeps", C.c_int),
                ("last_solve_ms", C.c_double)]


def bind(lib):
    L = lib.lib
    L.hs_set_options.argtypes = [iface.RCModelP, C.POINTER(Opts)]
    L.hs_get_options.argtypes = [iface.RCModelP, C.POINTER(Opts)]
    L.hs_get_stats.argtypes = [iface.RCModelP, C.POINTER(Stats)]
    L.hs_transient_run.argtypes = [iface.RCModelP, iface.c_double_p, C.c_int, C.c_double,
                                   iface.c_double_p, iface.c_double_p]
    L.hs_transient_run.restype = None
    L.hs_kernel_timer.argtypes = [C.c_char_p, C.c_int]
    L.hs_kernel_timer_read.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.hs_kernel_timer_read.restype = C.c_int
    L.hs_last_error.argtypes = [C.c_char_p, C.c_int]
    L.hs_last_error.restype = C.c_int


def check(lib):
    buf = C.create_string_buffer(512)
    if lib.lib.hs_last_error(buf, 512):
        raise RuntimeError(buf.value.decode())


def device_sync():
    hip = C.CDLL("libamdhip64.so")
    if hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def ptr(a):
    return a.ctypes.data_as(iface.c_double_p)


class Job:
    """The model, its block powers and the seeded power sequence."""

    def __init__(self, lib, n, intervals, seed=2):
        self.lib = lib
        self.case, pw = cases.stacked_case(n, seed=seed)
        self.run = cases.Run(lib, self.case, tempfile.mkdtemp(prefix="hsbench"))
        self.P = self.run.power_vector(pw)
        g = cases.rng(7)
        jit = g.uniform(0.8, 1.2, size=(intervals, self.run.n))
        self.seq = np.ascontiguousarray(self.P[None, :] * jit)
        self.intervals = intervals
        g_ = self.run.grid
        self.cells = g_.n_layers * g_.rows * g_.cols

    def options(self, precision, transient):
        o = Opts()
        self.lib.lib.hs_get_options(self.run.model, C.byref(o))
        o.precision, o.transient = precision, transient
        self.lib.lib.hs_set_options(self.run.model, C.byref(o))
        check(self.lib)

    def stats(self):
        s = Stats()
        self.lib.lib.hs_get_stats(self.run.model, C.byref(s))
        return s

    def once(self, keep_seq=False):
        """One job: steady solve, then the transient sequence from it."""
        L = self.lib.lib
        t0 = time.perf_counter()
        Ts = self.run.steady(self.P.copy())
        check(self.lib)
        t1 = time.perf_counter()
        out = np.zeros((self.intervals, self.run.n)) if keep_seq else None
        L.hs_transient_run(self.run.model, ptr(self.seq), self.intervals, DT, ptr(Ts),
                           ptr(out) if keep_seq else None)
        check(self.lib)
        t2 = time.perf_counter()
        return t2 - t0, t1 - t0, t2 - t1, Ts, out


def cpu_baseline(n, intervals_sample):
    """The reference library (compiled from its own sources) on one host core:
    compute_temp on the same 256 x 256 x 9 stack, a bounded sample of
    intervals from a uniform start, extrapolated per interval."""
    if not os.path.exists(REF_LIB):
        return None
    ref = iface.HotSpotLib(REF_LIB)
    case, pw = cases.stacked_case(n, seed=2)
    run = cases.Run(ref, case, tempfile.mkdtemp(prefix="hsref"))
    P = run.power_vector(pw)
    T0 = np.full(run.n, 333.15)
    seq = cases.jittered_powers(P, intervals_sample, seed=7)
    t0 = time.perf_counter()
    cases.run_transient(run, T0, seq, DT)
    dt = time.perf_counter() - t0
    g = run.grid
    cells = g.n_layers * g.rows * g.cols
    run.free()
    return dict(value=cells * intervals_sample / dt, unit="cell-updates/s", cores=1,
                kind="reference",
                sample=f"reference compute_temp, {intervals_sample} intervals of 3.333 us on the "
                       f"{n}x{n}x{g.n_layers} stack ({dt:.1f} s), single-threaded; the steady "
                       f"solve (1 of {1 + 10000} cell-updates per cell) is not sampled")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--intervals", type=int, default=10000)
    ap.add_argument("--cpu-intervals", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested).", file=sys.stderr)
        sys.exit(2)

    lib = iface.HotSpotLib(PROD_LIB)
    lib.lib.hs_set_fatal_mode(0)
    bind(lib)
    L = lib.lib

    log(f"building {args.grid}x{args.grid} 3D stack, {args.intervals} intervals")
    job = Job(lib, args.grid, args.intervals)
    job.options(1, 1)    # mixed precision, Taylor transient
    for w in range(args.warmup):
        tw, _, _, _, _ = job.once()
        log(f"warmup {w}: {tw:.3f} s")
    device_sync()
    times, steady_t, trans_t = [], [], []
    for k in range(args.steps):
        t, ts, tt, _, _ = job.once()
        times.append(t)
        steady_t.append(ts)
        trans_t.append(tt)
        log(f"step {k}: {t:.3f} s (steady {ts*1e3:.1f} ms, transient {tt:.3f} s)")
    device_sync()
    # roofline pass: the same job with every launch of the dominant kernel
    # bracketed by device events on the library's stream (plain launches
    # instead of graph replay while the timer is on)
    L.hs_kernel_timer(b"k_tay_term", 20000)
    t_roof, _, _, _, _ = job.once()
    device_sync()
    avg_ms, tot_ms = C.c_double(), C.c_double()
    nt = L.hs_kernel_timer_read(C.byref(avg_ms), C.byref(tot_ms))
    L.hs_kernel_timer(None, 0)
    log(f"roofline pass {t_roof:.3f} s: {nt} launches of k_tay_term, mean {avg_ms.value*1e3:.2f} us")
    st = job.stats()
    cells = job.cells
    updates = cells * (1 + args.intervals)
    t_job = float(np.mean(times))
    value = updates / t_job
    nvec = cells + (iface.EXTRA)
    bytes_per_launch = nvec * 16   # FP32: read term, read+write accumulator, write next term
    achieved = bytes_per_launch / (avg_ms.value * 1e-3) / 1e9 if nt else None
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_k_tay_term_3dstack256.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")

    # mixed-precision result of one job, kept for the comparison with FP64
    _, _, _, Ts_mixed, seq_mixed = job.once(keep_seq=True)

    fp64 = None
    if not args.no_fp64:
        job.options(0, 0)    # FP64 Jacobi-CG + the reference's adaptive RK4
        t64, ts64, tt64, Ts64, seq64 = job.once(keep_seq=True)
        st64 = job.stats()
        err = float(max(np.abs(seq64 - seq_mixed).max(), np.abs(Ts64 - Ts_mixed).max()))
        fp64 = dict(value=updates / t64, job_s=t64, steady_ms=ts64 * 1e3, transient_s=tt64,
                    rk4_calls_per_interval=st64.rk4_calls / args.intervals,
                    max_abs_diff_K_vs_mixed=err)
        log(f"fp64 job {t64:.3f} s (steady {ts64*1e3:.1f} ms), max |dT| mixed vs fp64 {err:.2e} K")

    cpu = None
    if not args.no_cpu:
        log("timing the reference library on the host")
        cpu = cpu_baseline(args.grid, args.cpu_intervals)

    res = {
        "metric": "cell-updates/s + p50/p99 solve latency vs plain FP64 GPU at <=0.01 K max err",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_job * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-sweeps/f64-residual",
        "data": "synthetic (seeded floorplans and powers)",
        "config": {"workload": f"3D stack {args.grid}x{args.grid}x9 (4 dies + 3 bonds + spreader + sink), "
                               f"256 blocks, steady + {args.intervals} x 3.333us intervals",
                   "cells": cells, "intervals": args.intervals, "precision": "mixed",
                   "steady": "JPCG FP32 inside FP64 refinement", "transient": "Taylor (exponential) FP32/FP64"},
        "latency_ms": {"job_p50": float(np.percentile(times, 50)) * 1e3,
                       "job_p99": float(np.percentile(times, 99)) * 1e3,
                       "steady_p50": float(np.percentile(steady_t, 50)) * 1e3,
                       "steady_p99": float(np.percentile(steady_t, 99)) * 1e3,
                       "interval_mean": float(np.mean(trans_t)) / args.intervals * 1e3},
        "roofline": {"bound": "hbm", "kernel": "k_tay_term<float>",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic, "launches_timed": nt,
                     "avg_launch_us": avg_ms.value * 1e3,
                     "bytes_per_launch": bytes_per_launch,
                     "note": "working set per vector 2.4 MB (FP32): cache-resident"},
        "solver": {"steady_iters": st.steady_iters, "taylor_terms_max": st.taylor_terms,
                   "taylor_substeps": st.taylor_substeps},
        "fp64_gpu": fp64,
        "speedup_vs_fp64_gpu": (value / fp64["value"]) if fp64 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
