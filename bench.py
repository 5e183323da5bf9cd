#!/usr/bin/env python3
"""Benchmark of the MI355X grid thermal solver on its headline workload.

Workload: jobs on a 3D-stacked chip on a 256 x 256 grid, 4 silicon dies
(0.1 mm, k = 100 W/mK) interleaved with 3 bond layers (10 um, k = 1.5 W/mK)
over the default spreader and sink (9 layers, 589 824 cells), 64 seeded
slicing blocks per die with power density lognormal(ln 0.5 W/mm^2, 0.6);
each job is a steady solve, then 10 000 transient intervals of 3.333 us with
+-20 % seeded per-interval power jitter, starting from the steady state's
block temperatures.  One "step" serves a batch of such jobs (different seeds)
through the batched engine, in the mixed-precision form (FP32 stencil sweeps
inside FP64 refinement / FP64 residuals; Taylor transient with FP32 terms and
FP64 state); inputs are admitted into device memory before the timed region.

metric: delivered cell-updates/s = jobs x cells x (1 steady + intervals) /
step time, with the p50/p99 job latency.  The same first job also runs alone
through the single-model API (its result must equal the served one bit for
bit) and in the plain FP64 GPU form (FP64 Jacobi-CG steady, the reference's
adaptive RK4 transient); the JSON line carries their rates, the speedup and
the largest block-temperature difference between the mixed and FP64 forms.

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
# The engine runs its set-up, steady, transient and backward-Euler work on
# ten streams.  HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware
# queues (four by default); streams that share a queue run in its order, so
# a steady solve's set-up would wait behind a transient chunk of another
# stream.  Sixteen queues give every stream its own; set before the first
# HIP call.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

from hotspot_amd import cases, engine, iface  # noqa: E402

PROD_LIB = os.path.join(ROOT, "hotspot_amd", "libhotspot_amd.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libhotspot_ref.so")
HBM_PEAK_GBS = 8000.0
DT = 3.333e-6
# Bound on the Taylor truncation error per interval (K).  The metric allows
# 0.01 K against plain FP64; 1e-6 K per interval over the 10 000 intervals of
# a job bounds the accumulated truncation error by 0.01 K even without the
# damping of the thermal system (the measured error, FP32 rounding included,
# is reported with the FP64 job, max_abs_diff_K_vs_mixed).  The library's
# own default is 1e-7.
TAYLOR_TOL = 1e-6


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
                ("last_solve_ms", C.c_double), ("be_iters", C.c_int)]


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
        self.rows = None
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
        """precision/transient as the engine's forms; the mixed form solves
        steady states with multigrid-preconditioned CG, like the engine."""
        o = Opts()
        self.lib.lib.hs_get_options(self.run.model, C.byref(o))
        o.precision, o.transient = precision, transient
        o.steady_solver = 1 if precision == 1 else 0
        o.taylor_tol = TAYLOR_TOL
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


def _cpu_worker(task):
    """One host process: the reference library (compiled from its own
    sources) runs compute_temp on the stack for k intervals from the given
    block temperatures; returns the seconds taken."""
    n, k, T0 = task
    ref = iface.HotSpotLib(REF_LIB)
    case, pw = cases.stacked_case(n, seed=2)
    run = cases.Run(ref, case, tempfile.mkdtemp(prefix="hsref"))
    P = run.power_vector(pw)
    seq = cases.jittered_powers(P, k, seed=7)
    t0 = time.perf_counter()
    cases.run_transient(run, T0, seq, DT)
    dt = time.perf_counter() - t0
    run.free()
    return dt


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(n, k, T0):
    """The reference library on the host: compute_temp on the same stack for
    a bounded sample of k intervals from the steady state's block
    temperatures (per-interval cost is flat, so the rate extrapolates), once
    in one process and then in one single-threaded process per available
    core at the same time (the reference has no threads; independent jobs
    are how a host would use its cores).  Processes are spawned fresh, so
    none touches the GPU."""
    if not os.path.exists(REF_LIB):
        return None
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    avail = len(os.sched_getaffinity(0))
    procs = max(1, min(avail, 16))
    with ctx.Pool(1) as pool:
        t1 = pool.map(_cpu_worker, [(n, k, T0)])[0]
    with ctx.Pool(procs) as pool:
        tp = pool.map(_cpu_worker, [(n, k, T0)] * procs)
    case, _ = cases.stacked_case(n, seed=2)
    cells = n * n * (len(case.lcf) + 2)
    one = cells * k / t1
    allc = cells * k * procs / max(tp)
    return dict(value=allc, unit="cell-updates/s", cores=procs, kind="reference",
                sample=f"reference compute_temp, {k} intervals of 3.333 us on the {n}x{n}x{len(case.lcf) + 2} "
                       f"stack from the steady state's block temperatures, in each of {procs} single-threaded "
                       f"processes at once ({max(tp):.1f} s); the steady solve (1 of 10001 cell-updates per cell) "
                       f"is not sampled",
                single_core=dict(value=one, seconds=t1),
                nproc=os.cpu_count(), cores_available=avail, cpu_model=cpu_model())


def served(lib, jobs, intervals, tick, keep="none", timer=None, form=None):
    """One batch of jobs through the engine (by default the mixed form: MG-PCG
    steady, FP32 Taylor transient): the jobs are submitted and admitted
    (inputs copied into device memory) before the timed region, which runs
    ticks until every result is back on the host.  Returns (timed seconds,
    admission seconds, job infos, block-temperature sequences: none, the
    first job's, or all)."""
    prec, trans = form or (engine.MIXED, engine.TAYLOR)
    eng = engine.Engine(lib, prec, trans)
    eng.set_options(precision=prec, steady_solver=1 if prec == engine.MIXED else 0, transient=trans,
                    taylor_tol=TAYLOR_TOL)
    eng.set_tick(tick, 128)
    ids = [eng.submit(j.run.model, j.rows, n_steps=intervals, dt=DT) for j in jobs]
    ta = time.perf_counter()
    if eng.admit() != 0:
        raise RuntimeError("the batch does not fit the engine's device-memory budget")
    device_sync()
    t0 = time.perf_counter()
    if timer:
        lib.lib.hs_kernel_timer(timer, 10 ** 7)
    eng.run()
    t1 = time.perf_counter()
    infos = [eng.info(i) for i in ids]
    seqs = None
    if keep == "first":
        seqs = [eng.result_seq(ids[0])]
    elif keep == "all":
        seqs = [eng.result_seq(i) for i in ids]
    for i in ids:
        eng.release(i)
    eng.close()
    return t1 - t0, t0 - ta, infos, seqs


def mode_many(lib, args):
    """Batched workload: uneven problems (grids drawn from 32^2..256^2, 2-4
    chip layers over spreader and sink, 8-128 slicing blocks, powers
    U(0.1, 10) W), each a steady solve then `--many-steps` intervals, all
    submitted together to the engine (mixed precision, Taylor); inputs
    admitted before the timed region."""
    g = cases.rng(41)
    jobs = []
    log(f"building {args.many} problems")
    for k in range(args.many):
        n = int(g.choice([32, 64, 128, 256]))
        cl = int(g.integers(2, 5))
        nb = int(g.choice([8, 16, 32, 64, 128]))
        case, pw = cases.chip_stack_case(n, seed=1000 + k, chip_layers=cl, n_blocks=nb, power=(0.1, 10.0))
        run = cases.Run(lib, case, tempfile.mkdtemp(prefix="hsmany"))
        P = run.power_vector(pw)
        seq = np.vstack([P[None, :], P[None, :] * cases.rng(2000 + k).uniform(0.8, 1.2, (args.many_steps, run.n))])
        gr = run.grid
        jobs.append((run, np.ascontiguousarray(seq), gr.n_layers * gr.rows * gr.cols))
    cells = sum(c for _, _, c in jobs)
    updates = sum(c * (1 + args.many_steps) for _, _, c in jobs)
    res = {}
    forms = [("mixed", engine.MIXED, engine.TAYLOR)] + ([] if args.no_fp64 else [("fp64", engine.FP64, engine.RK4)])
    for name, prec, trans in forms:
        eng = engine.Engine(lib, prec, trans)
        eng.set_tick(args.tick, 128)
        ids = [eng.submit(r.model, seq, n_steps=args.many_steps, dt=DT) for r, seq, _ in jobs]
        if eng.admit() != 0:
            raise RuntimeError("the batch does not fit the engine's device-memory budget")
        device_sync()
        t0 = time.perf_counter()
        eng.run()
        t = time.perf_counter() - t0
        lat = [eng.info(i).latency_s for i in ids]
        finals = [eng.poll(i) for i in ids]
        for i in ids:
            eng.release(i)
        eng.close()
        res[name] = dict(value=updates / t, seconds=t, latency_ms_p50=float(np.percentile(lat, 50)) * 1e3,
                         latency_ms_p99=float(np.percentile(lat, 99)) * 1e3, finals=finals)
        log(f"{name}: {len(jobs)} problems in {t:.2f} s, {updates / t:.3e} cell-updates/s")
    err = None
    if "fp64" in res:
        err = float(max(np.abs(a - b).max() for a, b in zip(res["mixed"]["finals"], res["fp64"]["finals"])))
    for v in res.values():
        v.pop("finals")
    for r, _, _ in jobs:
        r.free()
    return {"metric": "cell-updates/s (batched uneven problems, mixed precision)", "mode": "many",
            "value": res["mixed"]["value"], "unit": "cell-updates/s", "n_gpus": 1,
            "config": {"workload": f"{args.many} problems, grids 32^2-256^2, 4-6 layers, 8-128 blocks, "
                                   f"steady + {args.many_steps} intervals of 3.333us each",
                       "cells_total": cells, "problems": args.many},
            "forms": res, "max_abs_diff_K_mixed_vs_fp64": err,
            "speedup_vs_fp64": (res["mixed"]["value"] / res["fp64"]["value"]) if "fp64" in res else None}


def mode_serve(lib, args):
    """Served stream: seeded Poisson arrivals (rate --serve-rate) over
    --serve-seconds on a pool of resident models (grids 64^2..512^2, 4-8
    layers, 10 % microchannel stacks; 256-block floorplans with 8 hotspots
    at 10x the mean density 0.5 W/mm^2); 60 % steady solves, 40 % transient
    sequences of U{100..10000} intervals scaled by --serve-steps-scale.
    Requests are submitted when they arrive (wall clock) and the engine
    ticks continuously; latency is submission -> result."""
    g = cases.rng(77)
    pool = []
    log(f"building a pool of {args.serve_pool} resident models")
    for k in range(args.serve_pool):
        n = int(g.choice([64, 128, 256, 512]))
        if k % 10 == 9:
            d = tempfile.mkdtemp(prefix="hsmc")
            case, pw = cases.microchannel_case(n, 300 + k, d, die=16e-3, width_cells=2, pitch_cells=4)
            run = cases.Run(lib, case, d)
        else:
            cl = int(g.integers(2, 7))
            case, pw = cases.chip_stack_case(n, seed=500 + k, chip_layers=cl, n_blocks=256,
                                             power=("density", 0.5), hotspots=8)
            run = cases.Run(lib, case, tempfile.mkdtemp(prefix="hsserve"))
        gr = run.grid
        pool.append((run, run.power_vector(pw), gr.n_layers * gr.rows * gr.cols))
    # the request stream
    t, reqs = 0.0, []
    while True:
        t += g.exponential(1.0 / args.serve_rate)
        if t > args.serve_seconds:
            break
        m = int(g.integers(len(pool)))
        steps = 0 if g.random() < args.serve_steady_frac else max(1, int(round(int(g.integers(100, 10001)) * args.serve_steps_scale)))
        reqs.append((t, m, steps, pool[m][1] * g.uniform(0.8, 1.2, size=pool[m][1].shape)))
    updates = sum(pool[m][2] * (steps if steps else 1) for _, m, steps, _ in reqs)
    log(f"{len(reqs)} requests, {sum(1 for r in reqs if r[2])} transient, {updates:.3e} cell-updates")
    out = {}
    forms = [("mixed", engine.MIXED, engine.TAYLOR)] + ([] if args.no_fp64 else [("fp64", engine.FP64, engine.RK4)])
    for name, prec, trans in forms:
        eng = engine.Engine(lib, prec, trans)
        eng.set_tick(args.tick, 128)
        ids, k = [], 0
        t0 = time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            while k < len(reqs) and reqs[k][0] <= now:
                _, m, steps, P = reqs[k]
                ids.append(eng.submit(pool[m][0].model, P, n_steps=steps, dt=DT))
                k += 1
            active = eng.tick()
            if k == len(reqs) and active == 0:
                break
            if active == 0 and k < len(reqs):
                time.sleep(max(0.0, reqs[k][0] - (time.perf_counter() - t0)))
        wall = time.perf_counter() - t0
        est = eng.stats()
        infos = [eng.info(i) for i in ids]
        lat = np.array([x.latency_s for x in infos])
        steady_lat = np.array([x.latency_s for x, r in zip(infos, reqs) if not r[2]])
        out[name] = dict(value=updates / wall, wall_s=wall, p50_ms=float(np.percentile(lat, 50)) * 1e3,
                         p99_ms=float(np.percentile(lat, 99)) * 1e3,
                         steady_p50_ms=float(np.percentile(steady_lat, 50)) * 1e3 if len(steady_lat) else None,
                         steady_p99_ms=float(np.percentile(steady_lat, 99)) * 1e3 if len(steady_lat) else None,
                         wait_p50_ms=float(np.percentile([x.wait_s for x in infos], 50)) * 1e3,
                         setup_mean_ms=float(np.mean([x.setup_s for x in infos])) * 1e3,
                         resident_frac=float(np.mean([x.resident_model for x in infos])),
                         steady_p50_ms_by_cells={str(c): float(np.percentile(
                             [x.latency_s for x, r in zip(infos, reqs) if not r[2] and pool[r[1]][2] == c], 50)) * 1e3
                             for c in sorted({pool[r[1]][2] for r in reqs if not r[2]})},
                         engine={k: getattr(est, k) for k, _ in est._fields_})
        for i in ids:
            eng.release(i)
        eng.close()
        log(f"{name}: wall {wall:.2f} s, p50 {out[name]['p50_ms']:.0f} ms, p99 {out[name]['p99_ms']:.0f} ms")
    for r, _, _ in pool:
        r.free()
    return {"metric": "p50/p99 solve latency + cell-updates/s (served Poisson stream)", "mode": "serve",
            "value": out["mixed"]["value"], "unit": "cell-updates/s", "n_gpus": 1,
            "config": {"workload": f"Poisson {args.serve_rate}/s for {args.serve_seconds} s, pool of "
                                   f"{args.serve_pool} models 64^2-512^2 x 4-8 layers (10 % microchannel), "
                                   f"{100 * args.serve_steady_frac:.0f} % steady / {100 - 100 * args.serve_steady_frac:.0f} % "
                                   f"transient of U{{100..10000}} x {args.serve_steps_scale} intervals",
                       "requests": len(reqs), "cell_updates": updates},
            "forms": out,
            "speedup_vs_fp64": (out["mixed"]["value"] / out["fp64"]["value"]) if "fp64" in out else None}


def mode_large(lib, args):
    """Large steady solve: a --large-grid^2 x 8-layer stack with two
    microchannel layers (2-cell channels on a 4-cell pitch, 25.6 mm die,
    water at 300 K), 256 blocks with 8 hotspots; FP32 multigrid-preconditioned
    GMRES inside FP64 refinement to a 1e-12 relative FP64 residual."""
    d = tempfile.mkdtemp(prefix="hslarge")
    log(f"building the {args.large_grid}^2 two-channel stack")
    t0 = time.perf_counter()
    case, pw = cases.two_channel_stack_case(args.large_grid, 7, d)
    run = cases.Run(lib, case, d)
    log(f"model built in {time.perf_counter() - t0:.1f} s")
    P = run.power_vector(pw)
    o = Opts()
    L = lib.lib
    L.hs_get_options(run.model, C.byref(o))
    o.precision, o.steady_solver = 1, 1
    L.hs_set_options(run.model, C.byref(o))
    check(lib)
    run.steady(P.copy())    # warm-up: device set-up and multigrid hierarchy
    check(lib)
    times, st = [], None
    for k in range(args.steps):
        t = time.perf_counter()
        T = run.steady(P.copy())
        check(lib)
        times.append(time.perf_counter() - t)
        s_ = Stats()
        L.hs_get_stats(run.model, C.byref(s_))
        st = s_
        log(f"steady {k}: {times[-1]*1e3:.0f} ms, {st.steady_iters} GMRES iterations, "
            f"{st.steady_refinements} refinements, residual {st.steady_rel_resid:.1e}")
    g = run.grid
    cells = g.n_layers * g.rows * g.cols
    res = {"metric": "steady solve time (large two-microchannel stack)", "mode": "large",
           "value": cells / float(np.mean(times)), "unit": "cell-updates/s", "n_gpus": 1,
           "ms_per_solve": float(np.mean(times)) * 1e3, "device_ms": st.last_solve_ms,
           "gmres_iterations": st.steady_iters, "refinements": st.steady_refinements,
           "rel_residual_fp64": st.steady_rel_resid, "t_max_K": float(T.max()),
           "config": {"workload": f"{args.large_grid}x{args.large_grid}x8, two microchannel layers, 256 blocks, "
                                  f"8 hotspots", "cells": cells}}
    run.free()
    return res


def roofline_pass(lib, args, jobs, nvec):
    """The Horner sweep k_tay_term4 (one step h <- E_1 + s_j M h + t_j h per
    launch) timed by device events on the engine's stream: plain launches
    instead of graph replay while the timer is on, one transient stream;
    consecutive launches of the kernel share one pair of events around their
    run, so the mean duration per launch is the run's span over its launch
    count, gaps included.  Algorithmic bytes for one job: FP32 read of h,
    read of E_1, write of the new h = 12 B per node; 8 B when h is E_1 itself
    (a job's first step of a sub-step); 24 B for the last step, which adds
    the increment into the FP64 state.  A launch skips the jobs whose series
    needs fewer steps, so the achieved rate is the bytes of the steps the
    jobs took (counted on the device per job: taylor_node_bytes x nodes) over
    the summed duration of every launch.  The pass serves --roofline-batch
    jobs (the given ones first) so that one launch's working set (3 FP32
    vectors per job) is above the 256 MB Infinity Cache: the bytes are HBM
    bytes.  Returns (GB/s, mean ms per launch, total ms, launches, node
    bytes summed over the jobs)."""
    L = lib.lib
    kernel = "k_tay_term"
    rb = args.roofline_batch
    rjobs = jobs + [Job(lib, args.grid, args.roofline_intervals, seed=2 + b) for b in range(len(jobs), rb)]
    full_rows = [j.rows for j in jobs]
    for j in rjobs:
        j.rows = np.ascontiguousarray(np.vstack([j.P[None, :], j.seq[:args.roofline_intervals]]))
    served(lib, rjobs, args.roofline_intervals, args.tick)
    _, _, rinfo, _ = served(lib, rjobs, args.roofline_intervals, args.tick, timer=kernel.encode())
    device_sync()
    a, tot = C.c_double(), C.c_double()
    nt = L.hs_kernel_timer_read(C.byref(a), C.byref(tot))
    L.hs_kernel_timer(None, 0)
    avg_ms, tot_ms = a.value, tot.value
    node_bytes = float(sum(i.taylor_node_bytes for i in rinfo))
    achieved = None
    if nt:
        achieved = node_bytes * nvec / (tot_ms * 1e-3) / 1e9
        log(f"roofline pass ({rb} jobs): {nt} launches of {kernel}, mean {avg_ms*1e3:.2f} us, "
            f"{node_bytes * nvec / nt / 1e6:.1f} MB per launch on average, {achieved:.0f} GB/s")
    for j in rjobs[len(jobs):]:
        j.run.free()
    for j, r in zip(jobs, full_rows):
        j.rows = r
    return achieved, avg_ms, tot_ms, nt, node_bytes


def main():
    global TAYLOR_TOL
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["batch", "many", "serve", "large"], default="batch",
                    help="batch: the headline served batch (default); many: uneven batched problems; "
                         "serve: Poisson request stream; large: 1024^2 x 8 two-microchannel steady solve")
    ap.add_argument("--large-grid", type=int, default=1024)
    ap.add_argument("--many", type=int, default=512)
    ap.add_argument("--many-steps", type=int, default=500)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-pool", type=int, default=20)
    ap.add_argument("--serve-steps-scale", type=float, default=0.1)
    ap.add_argument("--serve-steady-frac", type=float, default=0.6, help="share of steady requests")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--intervals", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=16, help="jobs served together per step")
    ap.add_argument("--tick", type=int, default=100, help="transient intervals per engine tick")
    ap.add_argument("--cpu-intervals", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the served Poisson stream")
    ap.add_argument("--roofline-batch", type=int, default=48,
                    help="jobs of the roofline pass (one launch's working set above the Infinity Cache)")
    ap.add_argument("--roofline-intervals", type=int, default=300)
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the roofline pass (for kernel traces and counters) and print its figures")
    ap.add_argument("--taylor-tol", type=float, default=TAYLOR_TOL,
                    help="bound on the Taylor truncation error per interval (K)")
    args = ap.parse_args()
    TAYLOR_TOL = args.taylor_tol
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested).", file=sys.stderr)
        sys.exit(2)

    lib = iface.HotSpotLib(PROD_LIB)
    lib.lib.hs_set_fatal_mode(0)
    bind(lib)
    L = lib.lib
    if args.mode != "batch":
        res = {"many": mode_many, "serve": mode_serve, "large": mode_large}[args.mode](lib, args)
        print(json.dumps(res), flush=True)
        return

    if args.roofline_only:
        jobs = [Job(lib, args.grid, args.roofline_intervals, seed=2 + b) for b in range(args.roofline_batch)]
        nvec = jobs[0].cells + iface.EXTRA
        achieved, avg_ms, tot_ms, nt, node_bytes = roofline_pass(lib, args, jobs, nvec)
        print(json.dumps({"kernel": "k_tay_term4", "jobs": args.roofline_batch,
                          "intervals": args.roofline_intervals, "launches": nt,
                          "avg_launch_us": avg_ms * 1e3, "achieved_GBs": achieved,
                          "frac": achieved / HBM_PEAK_GBS if achieved else None,
                          "algorithmic_bytes_avg_per_launch": node_bytes * nvec / nt if nt else None}),
              flush=True)
        return

    log(f"building {args.batch} jobs: {args.grid}x{args.grid} 3D stack, {args.intervals} intervals")
    jobs = [Job(lib, args.grid, args.intervals, seed=2 + b) for b in range(args.batch)]
    for j in jobs:
        j.rows = np.ascontiguousarray(np.vstack([j.P[None, :], j.seq]))
    cells = jobs[0].cells
    updates_job = cells * (1 + args.intervals)
    updates_batch = args.batch * updates_job

    # ---- the headline: the batch through the engine in the mixed form ---------
    for w in range(args.warmup):
        tw, _, _, _ = served(lib, jobs, args.intervals, args.tick)
        log(f"warmup {w}: batch of {args.batch} in {tw:.3f} s")
    times, admit_t, lat = [], [], []
    for k in range(args.steps):
        t, ta, infos, _ = served(lib, jobs, args.intervals, args.tick)
        times.append(t)
        admit_t.append(ta)
        lat += [i.latency_s for i in infos]
        log(f"step {k}: batch of {args.batch} in {t:.3f} s (admission {ta*1e3:.0f} ms), "
            f"steady iters {[i.steady_iters for i in infos][:4]}...")
    t_step = float(np.mean(times))
    value = updates_batch / t_step

    # ---- where the speedup comes from: the same batch in other forms ----------
    # every job of the batch is kept in each form, so accuracy is stated over
    # the whole batch: mixed against FP64 (the metric's 0.01 K), and each
    # served job against the same job alone through the single-model API
    t_mix, _, _, seq_mix = served(lib, jobs, args.intervals, args.tick, keep="all")
    forms = {}
    fp64_batch_seqs = None
    if not args.no_fp64:
        for name, form in (("fp64_rk4", (engine.FP64, engine.RK4)), ("fp64_taylor", (engine.FP64, engine.TAYLOR))):
            tf, _, finfo, fseq = served(lib, jobs, args.intervals, args.tick, form=form,
                                       keep="all" if name == "fp64_rk4" else "none")
            forms[name] = dict(value=updates_batch / tf, seconds=tf,
                               rk4_calls_per_interval=float(np.mean([i.rk4_calls for i in finfo])) / args.intervals)
            if name == "fp64_rk4":
                fp64_batch_seqs = fseq
            log(f"batch in form {name}: {tf:.3f} s, {updates_batch / tf:.3e} cell-updates/s")
    forms["mixed_taylor"] = dict(value=value, seconds=t_step)

    # ---- the same jobs alone through the single-model API ---------------------
    single = None
    alone_equal = True
    for b, j in enumerate(jobs):
        j.options(1, 1)
        if b == 0:
            j.once()    # device set-up and term tiling outside the timing
        t1s, ts1, tt1, Ts_b, seq_b = j.once(keep_seq=True)
        alone_equal &= bool(np.array_equal(seq_b, seq_mix[b]))
        if b == 0:
            st = j.stats()
            single = dict(value=updates_job / t1s, job_s=t1s, steady_ms=ts1 * 1e3, transient_s=tt1,
                          steady_inner_iters=st.steady_iters, steady_refinements=st.steady_refinements,
                          taylor_terms_max=st.taylor_terms, taylor_substeps=st.taylor_substeps)
            log(f"single mixed job {t1s:.3f} s (steady {ts1*1e3:.1f} ms, transient {tt1:.3f} s)")
    single["served_equals_single_all_jobs"] = alone_equal
    log(f"every served job equals its single-model run bit for bit: {alone_equal}")

    # ---- the plain FP64 GPU form: one job at a time ---------------------------
    fp64 = None
    Ts64 = None
    if not args.no_fp64:
        j0 = jobs[0]
        j0.options(0, 0)    # FP64 Jacobi-CG + the reference's adaptive RK4
        t64, ts64, tt64, Ts64, seq64 = j0.once(keep_seq=True)
        st64 = j0.stats()
        err_batch = float(max(np.abs(a - b).max() for a, b in zip(seq_mix, fp64_batch_seqs)))
        fp64 = dict(value=updates_job / t64, job_s=t64, steady_ms=ts64 * 1e3, transient_s=tt64,
                    rk4_calls_per_interval=st64.rk4_calls / args.intervals,
                    max_abs_diff_K_vs_mixed=float(np.abs(seq64 - seq_mix[0]).max()),
                    max_abs_diff_K_mixed_vs_fp64_whole_batch=err_batch,
                    fp64_batch_equals_single=bool(np.array_equal(seq64, fp64_batch_seqs[0])))
        log(f"fp64 job {t64:.3f} s (steady {ts64*1e3:.1f} ms); max |dT| mixed vs fp64 over the batch "
            f"{err_batch:.2e} K")
    del seq_mix, fp64_batch_seqs

    # ---- roofline pass of the dominant kernel ----------------------------------
    kernel = "k_tay_term"
    nvec = cells + iface.EXTRA
    rb = args.roofline_batch
    achieved = avg_ms = nt = None
    node_bytes = tot_ms = 0.0
    if not args.no_roofline:
        achieved, avg_ms, tot_ms, nt, node_bytes = roofline_pass(lib, args, jobs, nvec)
    traffic = traffic_ratio = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{kernel}_3dstack{args.grid}_batch{rb}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        traffic, traffic_ratio = tj.get("bytes_per_launch"), tj.get("traffic_over_algorithmic")

    # ---- the served Poisson stream (requests arriving while others run) -------
    stream = None
    if not args.no_serve:
        stream = mode_serve(lib, args)

    # ---- the reference library on the host --------------------------------------
    cpu = None
    if not args.no_cpu:
        log("timing the reference library on the host")
        T0 = Ts64 if Ts64 is not None else jobs[0].run.steady(jobs[0].P.copy())
        cpu = cpu_baseline(args.grid, args.cpu_intervals, T0)

    lat_note = "submission -> result on the host, admission included"
    latency = {"batch_job_p50": float(np.percentile(lat, 50)) * 1e3,
               "batch_job_p99": float(np.percentile(lat, 99)) * 1e3,
               "batch_note": "the headline batch: all jobs submitted together; " + lat_note}
    if stream:
        for form_name, f in stream["forms"].items():
            latency[f"stream_{form_name}"] = {k: f[k] for k in ("p50_ms", "p99_ms", "steady_p50_ms", "steady_p99_ms")}
        latency["stream_note"] = ("served Poisson stream (" + stream["config"]["workload"] + "); " + lat_note)
    res = {
        "metric": "cell-updates/s + p50/p99 solve latency vs plain FP64 GPU at <=0.01 K max err",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-sweeps/f64-residual",
        "data": "synthetic (seeded floorplans and powers)",
        "config": {"workload": f"{args.batch} jobs served together, each a 3D stack {args.grid}x{args.grid}x9 "
                               f"(4 dies + 3 bonds + spreader + sink, 256 blocks): steady + "
                               f"{args.intervals} x 3.333us intervals",
                   "cells_per_job": cells, "intervals": args.intervals, "batch": args.batch,
                   "tick_intervals": args.tick, "precision": "mixed",
                   "steady": "multigrid-preconditioned CG, FP32 inside FP64 refinement",
                   "transient": "Taylor (exponential) FP32/FP64", "taylor_tol_K": TAYLOR_TOL},
        "latency_ms": latency,
        "admission_ms": float(np.mean(admit_t)) * 1e3,
        "value_with_admission": updates_batch / (t_step + float(np.mean(admit_t))),
        "roofline": {"bound": "hbm", "kernel": "k_tay_term4 (FP32 Horner step of the Taylor series, float4 tiles)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic, "launches_timed": nt,
                     "avg_launch_us": avg_ms * 1e3 if avg_ms else None,
                     "algorithmic_bytes_avg_per_launch": (node_bytes * nvec / nt) if nt else None,
                     "jobs_per_launch": rb, "intervals": args.roofline_intervals,
                     "working_set_per_launch_MB": rb * nvec * 12 / 1e6,
                     "traffic_over_algorithmic": traffic_ratio,
                     "note": f"{rb} jobs per launch: 3 FP32 vectors per job, {rb * nvec * 12 / 1e6:.0f} MB per "
                             f"launch, above the 256 MB Infinity Cache; traffic from rocprofv3 FETCH_SIZE and "
                             f"WRITE_SIZE of the same pass, scaled against a plain copy"},
        "forms_batch": forms,
        "speedup_split": ({"batching_fp64_rk4_over_single_fp64": forms["fp64_rk4"]["value"] / fp64["value"],
                           "integrator_taylor_over_rk4_fp64_batched": forms["fp64_taylor"]["value"] /
                           forms["fp64_rk4"]["value"],
                           "precision_mixed_over_fp64_taylor_batched": value / forms["fp64_taylor"]["value"]}
                          if fp64 and "fp64_rk4" in forms else None),
        "single_job": single,
        "fp64_gpu": fp64,
        "speedup_vs_fp64_gpu": (value / fp64["value"]) if fp64 else None,
        "served_stream": ({"value_mixed": stream["forms"]["mixed"]["value"],
                           "value_fp64": stream["forms"].get("fp64", {}).get("value"),
                           "speedup_vs_fp64": stream["speedup_vs_fp64"],
                           "requests": stream["config"]["requests"],
                           "forms": stream["forms"]} if stream else None),
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
