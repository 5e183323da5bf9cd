"""Every tiling of the FP32 Taylor sweep gives the same result, bit for bit.

The float4 Horner step runs over per-layer tiles or z-marching column tiles
(tile width 16..256, 4 or 8 rows per thread); the library times the
candidates on first use of a grid shape and keeps the fastest.  A cell's
arithmetic does not depend on the tiling, so each tiling pinned through
hs_pin_sweep_tiling must reproduce the block temperatures of every interval
and the final grid state of the per-layer 16-column tiling exactly: square
and non-square grids, 3D stacks of 9 layers and the 4-layer default stack
with its package ring, interior tiles and tiles at every edge.
"""
import ctypes as C

import numpy as np
import pytest

from hotspot_amd import cases, iface
from helpers import bind_extensions, check_error, grid_temp, make, ptr, set_options

pytestmark = pytest.mark.gpu

DT = 3.333e-6
STEPS = 6


@pytest.fixture(scope="module")
def lib(prod, gpu):
    L = bind_extensions(prod)
    L.lib.hs_transient_run.argtypes = [iface.RCModelP, iface.c_double_p, C.c_int, C.c_double,
                                       iface.c_double_p, iface.c_double_p]
    L.lib.hs_transient_run.restype = None
    L.lib.hs_pin_sweep_tiling.argtypes = [iface.RCModelP, C.c_int, C.c_int, C.c_int]
    L.lib.hs_pin_sweep_tiling.restype = C.c_int
    L.lib.hs_kernel_timer.argtypes = [C.c_char_p, C.c_int]
    L.lib.hs_kernel_timer.restype = None
    L.lib.hs_kernel_timer_read.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.lib.hs_kernel_timer_read.restype = C.c_int
    return L


def shapes():
    c1, p1 = cases.stacked_case(128, seed=21, n_dies=4, blocks_per_die=16)
    c2, p2 = cases.stacked_case(128, seed=22, n_dies=2, blocks_per_die=8, grid_cols=256)
    c3 = cases.default_stack_case(256, grid_cols=64)
    p3 = [cases.rng(23).uniform(0.5, 5, 16)]
    return [("stack-128x128x9", c1, p1), ("stack-128x256x5", c2, p2), ("default-256x64x4", c3, p3)]


def tilings(nc):
    out = []
    for shift in range(4, min(8, int(np.log2(nc))) + 1):
        for rpt in (4, 8):
            for z in (0, 1):
                out.append((shift, rpt, z))
    return out


def run_pinned(lib, case, pw, tiling):
    run = make(lib, case)
    try:
        assert lib.lib.hs_pin_sweep_tiling(run.model, *tiling) == 0, tiling
        check_error(lib)
        set_options(lib, run, precision=1, transient=1)
        P = run.power_vector(pw)
        seq = np.ascontiguousarray(cases.jittered_powers(P, STEPS, seed=9))
        init = np.full(run.n, 325.0) + cases.rng(8).uniform(0, 8, run.n)
        out = np.zeros_like(seq)
        lib.lib.hs_transient_run(run.model, ptr(seq), STEPS, DT, ptr(init), ptr(out))
        check_error(lib)
        return out, grid_temp(lib, run, 1)
    finally:
        lib.lib.hs_pin_sweep_tiling(run.model, 0, 0, 0)
        run.free()


@pytest.mark.parametrize("name,case,pw", shapes(), ids=[s[0] for s in shapes()])
def test_every_tiling_equals_per_layer_tiles(lib, name, case, pw):
    nc = case.config["grid_cols"]
    ref_blocks, ref_grid = run_pinned(lib, case, pw, (4, 4, 0))
    assert np.all(np.isfinite(ref_grid)) and ref_grid.max() > 325.0
    for t in tilings(nc):
        blocks, grid = run_pinned(lib, case, pw, t)
        assert np.array_equal(blocks, ref_blocks), (name, t, float(np.abs(blocks - ref_blocks).max()))
        assert np.array_equal(grid, ref_grid), (name, t, float(np.abs(grid - ref_grid).max()))


@pytest.mark.parametrize("name,case,pw", shapes(), ids=[s[0] for s in shapes()])
def test_float4_sweep_runs_and_plain_launches_equal_graphs(lib, name, case, pw):
    """The shapes above take the float4 sweep (its launches are counted by
    the kernel timer, which also turns graph replay off): the plain launches
    give the graph-replayed result."""
    ref_blocks, ref_grid = run_pinned(lib, case, pw, (5, 4, 1))
    lib.lib.hs_kernel_timer(b"k_tay_term", 1 << 20)
    try:
        blocks, grid = run_pinned(lib, case, pw, (5, 4, 1))
        a, tot = C.c_double(), C.c_double()
        n = lib.lib.hs_kernel_timer_read(C.byref(a), C.byref(tot))
    finally:
        lib.lib.hs_kernel_timer(None, 0)
    assert n >= STEPS, (name, n)
    assert np.array_equal(blocks, ref_blocks) and np.array_equal(grid, ref_grid), name


def test_pin_refuses_invalid_tilings(lib):
    c, _ = cases.stacked_case(128, seed=21, n_dies=1, blocks_per_die=4, grid_cols=64)
    run = make(lib, c)
    try:
        for bad in ((3, 4, 0), (7, 4, 0), (9, 4, 0), (5, 2, 0), (5, 4, 2)):
            assert lib.lib.hs_pin_sweep_tiling(run.model, *bad) == -1, bad
            buf = C.create_string_buffer(512)
            assert lib.lib.hs_last_error(buf, 512) and b"sweep tiling" in buf.value, bad
    finally:
        run.free()
