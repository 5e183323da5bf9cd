"""Parity of the HIP path (through the C ABI) with the oracle and, where it
was built, the reference library, on seeded inputs.

Tolerances: FP64 GPU path within 1e-4 K of the
converged reference operator; mixed-precision path within 1e-2 K.
"""
import tempfile

import ctypes as C
import math

import numpy as np
import scipy.sparse as sp
import pytest

import hs_oracle as O
from hotspot_amd import cases
from helpers import (bind_extensions, check_error, grid_temp, make, oracle_for, ptr,
                     set_options, stats)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-4
MIXED_TOL = 1e-2


def _cases():
    out = [("default-64", cases.default_stack_case(64), [cases.rng(1).uniform(0.5, 5, 16)]),
           ("secondary-16", cases.default_stack_case(16, model_secondary=1),
            [cases.rng(1).uniform(0.5, 5, 16)]),
           ("avg-nonaligned-16", cases.default_stack_case(16, nbx=3, nby=5, grid_map_mode="avg"),
            [cases.rng(5).uniform(0.5, 5, 15)])]
    c, pw = cases.stacked_case(32, seed=2, blocks_per_die=16)
    out.append(("3d-stack-32", c, pw))
    c, pw = cases.stacked_case(16, seed=4, n_dies=2, blocks_per_die=8)
    c.lcf[0].flp = [u + (1.5e6 + 1e5 * i, 0.01 * (1 + 0.3 * (i % 3))) for i, u in enumerate(c.lcf[0].flp)]
    c.detailed_3D = True
    out.append(("detailed-3d-16", c, pw))
    return out


CASES = _cases()


@pytest.fixture(scope="module")
def lib(prod, gpu):
    return bind_extensions(prod)


@pytest.mark.parametrize("name,case,pw", CASES, ids=[c[0] for c in CASES])
def test_steady_fp64_matches_oracle(lib, name, case, pw):
    run = make(lib, case)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T = run.steady(P)
    check_error(lib)
    To, Tg = om.steady(P)
    assert np.abs(T - To).max() <= FP64_TOL
    G = grid_temp(lib, run, 0)
    assert np.abs(G - Tg).max() <= FP64_TOL
    st = stats(lib, run)
    assert st.steady_rel_resid <= 1e-10
    run.free()


@pytest.mark.parametrize("name,case,pw", CASES, ids=[c[0] for c in CASES])
def test_steady_mixed_matches_oracle(lib, name, case, pw):
    run = make(lib, case)
    set_options(lib, run, precision=1)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T = run.steady(P)
    check_error(lib)
    To, _ = om.steady(P)
    assert np.abs(T - To).max() <= MIXED_TOL
    st = stats(lib, run)
    assert st.steady_rel_resid <= 1e-10 and st.steady_refinements >= 1
    run.free()


def _mc_case():
    return cases.microchannel_case(16, 5, tempfile.mkdtemp())


# grid -> block in the min and max modes (temperature_grid.c:1932-2015) on a
# floorplan whose block edges fall inside cells
MAP_MODE_CASES = [(f"{mode}-nonaligned-16", cases.default_stack_case(16, nbx=3, nby=5, grid_map_mode=mode),
                   [cases.rng(6).uniform(0.5, 5, 15)]) for mode in ("min", "max")]
SLOPE_CASES = CASES + MAP_MODE_CASES + [("microchannel-16",) + _mc_case()]


@pytest.mark.parametrize("name,case,pw", SLOPE_CASES, ids=[c[0] for c in SLOPE_CASES])
def test_slope_and_maps_match_oracle(lib, name, case, pw):
    run = make(lib, case)
    om = oracle_for(run)
    P = run.power_vector(pw)
    # block -> grid power and temperature
    Pg = np.zeros(om.N)
    lib.lib.hs_block_to_grid(run.model, ptr(P), ptr(Pg), 0)
    check_error(lib)
    Po = om.b2g(P, power=True)
    Po[om.ncell:] = 0.0
    Pg[om.ncell:] = 0.0
    assert np.abs(Pg - Po).max() <= 1e-12 * max(1.0, np.abs(Po).max())
    Tb = 318.0 + 30 * cases.rng(11).random(run.n)
    Tg = np.zeros(om.N)
    lib.lib.hs_block_to_grid(run.model, ptr(Tb), ptr(Tg), 1)
    assert np.abs(Tg - om.b2g(Tb, power=False)).max() <= 1e-12
    # grid -> block
    y = 318.15 + 40 * cases.rng(3).random(om.N)
    back = np.zeros(run.n)
    lib.lib.hs_grid_to_block(run.model, ptr(y), ptr(back))
    assert np.abs(back - om.g2b(y)).max() <= 1e-12
    # operator: f(T) = (P + b - G T)/C
    s = np.zeros(om.N)
    lib.lib.hs_grid_slope(run.model, ptr(P), ptr(y), ptr(s))
    check_error(lib)
    so = om.slope(y, om.power_grid(P))
    assert np.abs(s - so).max() <= 1e-10 * np.abs(so).max()
    run.free()


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,case,pw", CASES, ids=[c[0] for c in CASES])
def test_steady_multigrid_matches_oracle(lib, name, case, pw, prec):
    """Multigrid-preconditioned FP32 CG inside FP64 refinement (either
    precision option) reaches the same 1e-12 FP64 residual: within the FP64
    bar of the oracle, and with far fewer inner iterations than Jacobi."""
    run = make(lib, case)
    set_options(lib, run, precision=prec, steady_solver=1)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T = run.steady(P.copy())
    check_error(lib)
    To, Tg = om.steady(P)
    assert np.abs(T - To).max() <= FP64_TOL
    assert np.abs(grid_temp(lib, run, 0) - Tg).max() <= FP64_TOL
    st = stats(lib, run)
    assert st.steady_rel_resid <= 1e-10
    mg_iters = st.steady_iters
    set_options(lib, run, precision=1, steady_solver=0)
    run.steady(P.copy())
    assert mg_iters < stats(lib, run).steady_iters
    run.free()


@pytest.mark.parametrize("solver", [0, 1], ids=["jacobi", "multigrid"])
@pytest.mark.parametrize("prec,tol", [(0, FP64_TOL), (1, MIXED_TOL)], ids=["fp64", "mixed"])
def test_microchannel_steady_gmres(lib, prec, tol, solver):
    """Advection makes G non-symmetric: with either solver option the steady
    solve runs restarted GMRES cycles inside FP64 refinement, preconditioned
    by the multigrid V-cycle of the upwind advection, and is checked against
    the oracle's direct sparse solve of the same (central) operator."""
    case, pw = _mc_case()
    run = make(lib, case)
    set_options(lib, run, precision=prec, steady_solver=solver)
    om = oracle_for(run)
    assert not om.symmetric
    P = run.power_vector(pw)
    T = run.steady(P)
    check_error(lib)
    To, Tg = om.steady(P)
    assert np.abs(T - To).max() <= tol
    assert np.abs(grid_temp(lib, run, 0) - Tg).max() <= tol
    st = stats(lib, run)
    assert st.steady_rel_resid <= 1e-10 and st.steady_iters > 0
    run.free()


@pytest.mark.parametrize("name,case,pw", [CASES[0], CASES[1], CASES[3], ("microchannel-16",) + _mc_case()],
                         ids=["default-64", "secondary-16", "3d-stack-32", "microchannel-16"])
def test_transient_rk4_matches_oracle(lib, name, case, pw):
    run = make(lib, case)
    om = oracle_for(run)
    P = run.power_vector(pw)
    if case.use_microchannels:
        T0 = np.full(run.n, 320.0)
    else:
        T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 10, seed=7)
    tr = cases.run_transient(run, T0, seq, 3.333e-6)
    check_error(lib)
    Tg = om.b2g(T0, power=False)
    for k, p in enumerate(seq):
        Tg, _ = om.advance(Tg, om.power_grid(p), 3.333e-6)
        assert np.abs(tr[k] - om.g2b(Tg)).max() <= 1e-8
    assert stats(lib, run).rk4_calls >= 1
    run.free()


# The Taylor integrator sums the exponential series to a 1e-7 K tail, so it is
# checked against the exact solution of each interval's linear ODE (the
# oracle's expm); the reference's adaptive RK4 itself sits up to ~2e-4 K from
# that solution when the first interval starts from block temperatures
# mapped onto the grid (piecewise constant per block, fast modes excited).
@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (1, 1e-4)], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,case,pw", [CASES[0], CASES[3]], ids=["default-64", "3d-stack-32"])
def test_transient_taylor_matches_exact(lib, name, case, pw, prec, tol):
    run = make(lib, case)
    set_options(lib, run, transient=1, precision=prec)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 10, seed=7)
    tr = cases.run_transient(run, T0, seq, 3.333e-6)
    check_error(lib)
    Tg = om.b2g(T0, power=False)
    worst = 0.0
    for k, p in enumerate(seq):
        Tg = om.advance_exact(Tg, om.power_grid(p), 3.333e-6)
        worst = max(worst, np.abs(tr[k] - om.g2b(Tg)).max())
    assert worst <= tol
    run.free()


# Long intervals take many sub-steps of rho tau <= 4, each with the series
# expanded about the centre of the spectrum (shift rho tau / 2) for
# conduction models; a microchannel model (coolant advection) keeps the
# unshifted series.  Both against the exact solution of each interval.
def _microchannel16():
    import tempfile
    return cases.microchannel_case(16, 5, tempfile.mkdtemp())


@pytest.mark.parametrize("dt", [1e-4, 1e-2], ids=["100us", "10ms"])
@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (1, 1e-4)], ids=["fp64", "mixed"])
@pytest.mark.parametrize("which", ["default-64", "secondary-16", "3d-stack-32", "detailed-3d-16", "microchannel-16"])
def test_transient_taylor_long_intervals_match_exact(lib, which, prec, tol, dt):
    if which == "microchannel-16":
        case, pw = _microchannel16()
    else:
        _, case, pw = next(c for c in CASES if c[0] == which)
    run = make(lib, case)
    set_options(lib, run, transient=1, precision=prec)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 3, seed=11)
    tr = cases.run_transient(run, T0, seq, dt)
    check_error(lib)
    st = stats(lib, run)
    # sub-steps keep rho tau <= 4, rho the Gershgorin bound on ||C^-1 G||
    A = abs(sp.diags(1.0 / om.Cvec) @ om.Gt)
    rho = float(np.asarray(A.sum(axis=1)).max())
    want = math.ceil(rho * dt / 4.0)
    assert abs(st.taylor_substeps - want) <= 1, (st.taylor_substeps, want)
    if want > 2:
        assert st.taylor_substeps > 1
    Tg = om.b2g(T0, power=False)
    worst = 0.0
    for k, p in enumerate(seq):
        Tg = om.advance_exact(Tg, om.power_grid(p), dt)
        worst = max(worst, np.abs(tr[k] - om.g2b(Tg)).max())
    assert worst <= tol, worst
    run.free()


def test_reference_steady_within_its_own_stopping_error(lib, ref):
    """The stock reference output stops its Gauss-Seidel sweeps at a 1e-6 K
    per-sweep change, which leaves it ~1e-3 K from the converged operator;
    the GPU result must sit within that band of the stock output."""
    case = cases.default_stack_case(64)
    pw = [cases.rng(1).uniform(0.5, 5, 16)]
    r = make(ref, case)
    p = make(lib, case)
    P = r.power_vector(pw)
    Tr = r.steady(P)
    Tp = p.steady(P.copy())
    assert np.abs(Tr - Tp).max() <= 5e-3
    r.free()
    p.free()


def test_reference_transient_trajectory(lib, ref):
    """Same CLI-style transient (steady start, jittered powers) through both
    libraries: the reference's RK4 on its stock steady state vs ours."""
    case = cases.default_stack_case(32)
    pw = [cases.rng(1).uniform(0.5, 5, 16)]
    r = make(ref, case)
    p = make(lib, case)
    P = r.power_vector(pw)
    T0 = r.steady(P)
    seq = cases.jittered_powers(P, 10, seed=7)
    tr = cases.run_transient(r, T0, seq, 3.333e-6)
    tp = cases.run_transient(p, T0, seq, 3.333e-6)
    assert np.abs(tr - tp).max() <= 1e-8
    r.free()
    p.free()


def test_microchannel_steady_multigrid_at_scale(lib):
    """A 64 x 64 microchannel stack (2-cell channels on a 4-cell pitch): the
    GMRES preconditioned by the upwind multigrid V-cycle converges in a few
    hundred iterations, to the oracle's direct solve."""
    d = tempfile.mkdtemp()
    case, pw = cases.microchannel_case(64, 3, d, die=16e-3, width_cells=2, pitch_cells=4)
    run = cases.Run(lib, case, d)
    set_options(lib, run, precision=1, steady_solver=1)
    P = run.power_vector(pw)
    T = run.steady(P.copy())
    check_error(lib)
    st = stats(lib, run)
    om = oracle_for(run)
    To, _ = om.steady(P)
    assert np.abs(T - To).max() <= FP64_TOL
    assert st.steady_rel_resid <= 1e-10 and st.steady_iters < 1000
    run.free()


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_two_channel_stack_128_default_solver(lib, prec):
    """The two-microchannel stack at 128 x 128 (cell Peclet number ~10^3 in
    the channels) with the default solver option: Jacobi-preconditioned GMRES
    stalls here (relative residual 1e-3 after 64 000 iterations); the upwind
    multigrid preconditioner converges in a few hundred, to the oracle's
    direct solve."""
    d = tempfile.mkdtemp()
    case, pw = cases.two_channel_stack_case(128, 7, d)
    run = cases.Run(lib, case, d)
    set_options(lib, run, precision=prec, steady_solver=0)
    P = run.power_vector(pw)
    T = run.steady(P.copy())
    check_error(lib)
    st = stats(lib, run)
    assert st.steady_rel_resid <= 1e-11 and st.steady_iters < 2000
    To, _ = oracle_for(run).steady(P)
    assert np.abs(T - To).max() <= (FP64_TOL if prec == 0 else MIXED_TOL)
    run.free()


def test_unconverged_steady_solve_is_an_error(lib):
    """A solve stopped by its iteration limit far above the tolerance reports
    an error instead of returning the unconverged state."""
    name, case, pw = CASES[0]
    run = make(lib, case)
    set_options(lib, run, max_iters=2)
    run.steady(run.power_vector(pw))
    buf = C.create_string_buffer(512)
    assert lib.lib.hs_last_error(buf, 512) == 1
    assert "did not converge" in buf.value.decode()
    run.free()


def test_microchannel_backward_euler_at_scale(lib):
    """Backward Euler on the 64 x 64 microchannel stack, 10 ms intervals
    (one implicit step each, FP64 GMRES), against the oracle's direct solve
    of the same step."""
    d = tempfile.mkdtemp()
    case, pw = cases.microchannel_case(64, 3, d, die=16e-3, width_cells=2, pitch_cells=4)
    run = cases.Run(lib, case, d)
    set_options(lib, run, transient=2)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 3, seed=4)
    tr = cases.run_transient(run, T0, seq, 1e-2)
    check_error(lib)
    Tg = om.b2g(T0, power=False)
    for k, p in enumerate(seq):
        Tg = om.advance_be(Tg, om.power_grid(p), 1e-2)
        assert np.abs(tr[k] - om.g2b(Tg)).max() <= 1e-5
    assert stats(lib, run).be_iters > 0
    run.free()


def test_two_channel_stack_multigrid(lib):
    """Device Si / Si / channels / Si / channels / Si cap over the package
    (the large steady configuration's stack) at 64 x 64: FP32 multigrid-
    preconditioned GMRES inside FP64 refinement against the oracle's direct
    solve."""
    d = tempfile.mkdtemp()
    case, pw = cases.two_channel_stack_case(64, 7, d)
    run = cases.Run(lib, case, d)
    set_options(lib, run, precision=1, steady_solver=1)
    P = run.power_vector(pw)
    T = run.steady(P.copy())
    check_error(lib)
    st = stats(lib, run)
    om = oracle_for(run)
    To, Tg = om.steady(P)
    assert np.abs(T - To).max() <= FP64_TOL
    assert np.abs(grid_temp(lib, run, 0) - Tg).max() <= FP64_TOL
    assert st.steady_rel_resid <= 1e-10
    run.free()


# Rectangular grids (sizes are powers of two, as the reference requires):
# 512 x 16 and 256 x 32 take float4 Taylor sweeps with 16- and 32-column
# tiles; 64 x 256 has four column tiles per row and scalar sweeps.
RECT = [("rect-512x16", cases.default_stack_case(64, grid_rows=512, grid_cols=16),
         [cases.rng(3).uniform(0.5, 5, 16)]),
        ("rect-256x32", cases.default_stack_case(64, grid_rows=256, grid_cols=32),
         [cases.rng(4).uniform(0.5, 5, 16)]),
        ("rect-64x256", cases.default_stack_case(64, grid_rows=64, grid_cols=256),
         [cases.rng(6).uniform(0.5, 5, 16)])]


@pytest.mark.parametrize("name,case,pw", RECT, ids=[c[0] for c in RECT])
def test_rectangular_grid_steady_multigrid(lib, name, case, pw):
    run = make(lib, case)
    set_options(lib, run, precision=1, steady_solver=1)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T = run.steady(P.copy())
    check_error(lib)
    To, Tg = om.steady(P)
    assert np.abs(T - To).max() <= FP64_TOL
    assert np.abs(grid_temp(lib, run, 0) - Tg).max() <= FP64_TOL
    assert stats(lib, run).steady_rel_resid <= 1e-10
    run.free()


@pytest.mark.parametrize("name,case,pw", RECT, ids=[c[0] for c in RECT])
def test_rectangular_grid_taylor_matches_exact(lib, name, case, pw):
    run = make(lib, case)
    set_options(lib, run, transient=1, precision=1)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 5, seed=9)
    tr = cases.run_transient(run, T0, seq, 3.333e-6)
    check_error(lib)
    Tg = om.b2g(T0, power=False)
    worst = 0.0
    for k, p in enumerate(seq):
        Tg = om.advance_exact(Tg, om.power_grid(p), 3.333e-6)
        worst = max(worst, np.abs(tr[k] - om.g2b(Tg)).max())
    assert worst <= 1e-4
    run.free()


# Backward Euler (the reference's SuperLU-build integrator): one implicit
# step per interval, against the oracle's restatement of the same step.  The
# CG stops at a 1e-14 relative residual of C T/h + P + b.  At 3.3 us the
# system mixes rows scaled by C/h with the capacitance-free rows of the
# secondary package nodes, and the CG and the direct solve then differ by up
# to ~2e-6 K there; the bar is 1e-5 K, ten times under the FP64 bar of
# 1e-4 K.  Short (3.333 us) and long (10 ms) intervals.  A microchannel
# model's system is non-symmetric (coolant advection) and is solved by FP64
# GMRES from the previous state, to the same tolerance.
@pytest.mark.parametrize("dt", [3.333e-6, 1e-2], ids=["3.3us", "10ms"])
@pytest.mark.parametrize("name,case,pw", [CASES[0], CASES[1], CASES[3], CASES[4], ("microchannel-16",) + _mc_case()],
                         ids=["default-64", "secondary-16", "3d-stack-32", "detailed-3d-16", "microchannel-16"])
def test_transient_backward_euler_matches_oracle(lib, name, case, pw, dt):
    run = make(lib, case)
    set_options(lib, run, transient=2)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 4, seed=5)
    tr = cases.run_transient(run, T0, seq, dt)
    check_error(lib)
    Tg = om.b2g(T0, power=False)
    for k, p in enumerate(seq):
        Tg = om.advance_be(Tg, om.power_grid(p), dt)
        assert np.abs(tr[k] - om.g2b(Tg)).max() <= 1e-5
    assert stats(lib, run).be_iters > 0
    run.free()
