"""Grids whose sizes are not powers of two, through the C ABI.

The reference's SuperLU build takes any grid size; its iterative build asks
for powers of two (temperature_grid.c:901-905) because of its multigrid.
Here every kernel covers partial tiles: the per-layer and column tiles mask
the cells past the grid's rows and columns, multigrid coarsens while both
sizes stay even (a grid of odd size keeps one level), and the float4 Taylor
sweeps take a problem only when its rows hold whole tiles (partial row tiles
are masked).  The same tolerances as the power-of-two parity tests, against
the oracle on the same inputs:
  steady FP64 and multigrid-preconditioned: 1e-4 K of the direct solve;
  operator and maps: 1e-10 relative;
  RK4: 1e-8 K of the oracle's restatement of the reference's RK4;
  Taylor (mixed): 1e-4 K of the exact solution of each interval;
  backward Euler: 1e-5 K of the oracle's implicit step.
Example 5 of the reference (a 39 x 39 microchannel stack, which its SuperLU
build runs) goes through the command line: steady outputs against the
oracle's direct solve, the transient with the SuperLU build's integrator
(-transient_solver be) against the oracle's backward-Euler steps.
"""
import os
import shutil
import subprocess
import tempfile

import ctypes as C
import numpy as np
import pytest

import hs_oracle as O
from hotspot_amd import cases, engine, iface
from helpers import bind_extensions, check_error, grid_temp, make, oracle_for, ptr, set_options, stats

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FP64_TOL = 1e-4
DT = 3.333e-6


def _cases():
    out = [("default-39", cases.default_stack_case(39), [cases.rng(1).uniform(0.5, 5, 16)]),
           ("default-48x80", cases.default_stack_case(64, grid_rows=48, grid_cols=80),
            [cases.rng(2).uniform(0.5, 5, 16)]),
           # float4 Taylor sweeps (64-column tiles) over 192 rows: whole
           # column tiles, rows of three tile heights
           ("default-192x64", cases.default_stack_case(64, grid_rows=192, grid_cols=64),
            [cases.rng(3).uniform(0.5, 5, 16)]),
           ("secondary-20", cases.default_stack_case(20, model_secondary=1), [cases.rng(4).uniform(0.5, 5, 16)])]
    c, pw = cases.stacked_case(40, seed=2, blocks_per_die=16)
    out.append(("3d-stack-40", c, pw))
    c, pw = cases.stacked_case(24, seed=4, n_dies=2, blocks_per_die=8)
    c.lcf[0].flp = [u + (1.5e6 + 1e5 * i, 0.01 * (1 + 0.3 * (i % 3))) for i, u in enumerate(c.lcf[0].flp)]
    c.detailed_3D = True
    out.append(("detailed-3d-24", c, pw))
    return out


CASES = _cases()
IDS = [c[0] for c in CASES]


def _mc39():
    return cases.microchannel_case(39, 5, tempfile.mkdtemp(prefix="hsnpot"))


@pytest.fixture(scope="module")
def lib(prod, gpu):
    return bind_extensions(prod)


@pytest.mark.parametrize("name,case,pw", CASES + [("microchannel-39",) + _mc39()], ids=IDS + ["microchannel-39"])
def test_npot_slope_maps_and_steady(lib, name, case, pw):
    print("npot slope/maps/steady:", name, flush=True)
    run = make(lib, case)
    om = oracle_for(run)
    assert (om.nr, om.nc) == (run.grid.rows, run.grid.cols)
    P = run.power_vector(pw)
    Pg = np.zeros(om.N)
    lib.lib.hs_block_to_grid(run.model, ptr(P), ptr(Pg), 0)
    check_error(lib)
    Po = om.b2g(P, power=True)
    Po[om.ncell:] = 0.0
    Pg[om.ncell:] = 0.0
    assert np.abs(Pg - Po).max() <= 1e-12 * max(1.0, np.abs(Po).max())
    y = 318.15 + 40 * cases.rng(3).random(om.N)
    back = np.zeros(run.n)
    lib.lib.hs_grid_to_block(run.model, ptr(y), ptr(back))
    assert np.abs(back - om.g2b(y)).max() <= 1e-12
    s = np.zeros(om.N)
    lib.lib.hs_grid_slope(run.model, ptr(P), ptr(y), ptr(s))
    check_error(lib)
    so = om.slope(y, om.power_grid(P))
    assert np.abs(s - so).max() <= 1e-10 * np.abs(so).max()
    To, Tg = om.steady(P)
    # FP64 Jacobi (symmetric) / Jacobi-GMRES (advection); then multigrid-
    # preconditioned FP32 inside FP64 refinement
    for prec, solver in ((0, 0), (1, 1)):
        set_options(lib, run, precision=prec, steady_solver=solver)
        T = run.steady(P.copy())
        check_error(lib)
        assert np.abs(T - To).max() <= FP64_TOL, (prec, solver)
        assert np.abs(grid_temp(lib, run, 0) - Tg).max() <= FP64_TOL, (prec, solver)
        assert stats(lib, run).steady_rel_resid <= 1e-10
    run.free()


@pytest.mark.parametrize("name,case,pw", [CASES[0], CASES[1], CASES[3], CASES[5]],
                         ids=[IDS[0], IDS[1], IDS[3], IDS[5]])
def test_npot_transient_rk4_matches_oracle(lib, name, case, pw):
    print("npot rk4:", name, flush=True)
    run = make(lib, case)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 5, seed=7)
    tr = cases.run_transient(run, T0, seq, DT)
    check_error(lib)
    Tg = om.b2g(T0, power=False)
    for k, p in enumerate(seq):
        Tg, _ = om.advance(Tg, om.power_grid(p), DT)
        assert np.abs(tr[k] - om.g2b(Tg)).max() <= 1e-8
    run.free()


@pytest.mark.parametrize("name,case,pw", CASES, ids=IDS)
def test_npot_transient_taylor_matches_exact(lib, name, case, pw):
    print("npot taylor:", name, flush=True)
    run = make(lib, case)
    set_options(lib, run, transient=1, precision=1)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 4, seed=9)
    tr = cases.run_transient(run, T0, seq, DT)
    check_error(lib)
    Tg = om.b2g(T0, power=False)
    worst = 0.0
    for k, p in enumerate(seq):
        Tg = om.advance_exact(Tg, om.power_grid(p), DT)
        worst = max(worst, float(np.abs(tr[k] - om.g2b(Tg)).max()))
    assert worst <= 1e-4
    run.free()


@pytest.mark.parametrize("name,case,pw", [CASES[0], CASES[3], ("microchannel-39",) + _mc39()],
                         ids=[IDS[0], IDS[3], "microchannel-39"])
def test_npot_backward_euler_matches_oracle(lib, name, case, pw):
    print("npot backward euler:", name, flush=True)
    run = make(lib, case)
    set_options(lib, run, transient=2)
    om = oracle_for(run)
    P = run.power_vector(pw)
    T0, _ = om.steady(P)
    seq = cases.jittered_powers(P, 3, seed=5)
    tr = cases.run_transient(run, T0, seq, 1e-3)
    check_error(lib)
    Tg = om.b2g(T0, power=False)
    for k, p in enumerate(seq):
        Tg = om.advance_be(Tg, om.power_grid(p), 1e-3)
        assert np.abs(tr[k] - om.g2b(Tg)).max() <= 1e-5
    run.free()


@pytest.mark.parametrize("form", [(engine.FP64, engine.RK4), (engine.MIXED, engine.TAYLOR)],
                         ids=["fp64-rk4", "mixed-taylor"])
def test_npot_engine_alone_equals_in_mix(lib, form):
    """Jobs on grids of any size share the engine's launches with power-of-
    two jobs; each result equals its result alone, bit for bit."""
    specs = [CASES[0], CASES[2], CASES[4], ("default-64", cases.default_stack_case(64),
                                            [cases.rng(9).uniform(0.5, 5, 16)])]
    runs = [make(lib, c) for _, c, _ in specs]
    rows = []
    for r, (_, _, pw) in zip(runs, specs):
        P = r.power_vector(pw)
        rows.append(np.vstack([P[None, :], np.array(cases.jittered_powers(P, 6, seed=3))]))
    print("npot engine mix:", [s[0] for s in specs], flush=True)
    eng = engine.Engine(lib, *form)
    ids = [eng.submit(r.model, rw, n_steps=6, dt=DT) for r, rw in zip(runs, rows)]
    eng.run()
    mix = [(eng.poll(i), eng.result_seq(i)) for i in ids]
    eng.close()
    for r, rw, (fin, seq), (name, _, _) in zip(runs, rows, mix, specs):
        e1 = engine.Engine(lib, *form)
        i = e1.submit(r.model, rw, n_steps=6, dt=DT)
        e1.run()
        assert np.array_equal(e1.poll(i), fin), name
        assert np.array_equal(e1.result_seq(i), seq), name
        e1.close()
    for r in runs:
        r.free()


# ---- the reference's example 5 through the command line ----------------------
EX5 = os.path.join(ROOT, "tests", "golden", "cli", "example5_mc")
EXE = os.path.join(ROOT, "hotspot_amd", "hotspot")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libhotspot_ref.so")


def _run_cli(args, cwd):
    print("cli:", " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]


def _read_block_file(path):
    out = {}
    for line in open(path):
        t = line.split()
        if len(t) == 2:
            out[t[0]] = float(t[1])
    return out


def _example5_oracle(d):
    """The oracle on example 5's files; the configuration as the reference
    library (compiled from its sources) parses it, with the materials file."""
    if not os.path.exists(REF_LIB):
        pytest.skip("reference library not built (oracle/_ref)")
    ref = iface.HotSpotLib(REF_LIB)
    t = (iface.StrPair * 512)()
    n = ref.lib.read_str_pairs(t, 512, os.path.join(d, "example.config").encode())
    ml = iface.MaterialsList()
    ref.lib.default_materials(C.byref(ml))
    ref.lib.materials_add_from_file(C.byref(ml), os.path.join(d, "example.materials").encode())
    tc = ref.lib.default_thermal_config()
    ref.lib.thermal_config_add_from_strs(C.byref(tc), C.byref(ml), t, n)
    mc = ref.lib.default_microchannel_config()
    ref.lib.microchannel_config_add_from_strs(C.byref(mc), C.byref(ml), t, n)
    cfg = {k: getattr(tc, k) for k in O.DEFAULT_CONFIG if hasattr(tc, k) and k not in
           ("grid_map_mode", "grid_layer_file")}
    cfg["grid_map_mode"] = tc.grid_map_mode.decode()
    mcd = {k: getattr(mc, k) for k in O.DEFAULT_MICROCHANNEL}
    mats = O.read_materials(os.path.join(d, "example.materials"))
    om = O.OracleModel(cfg, lcf=os.path.join(d, "example.lcf"), detailed_3D=True, use_microchannels=True,
                       microchannel=mcd, materials=mats)
    return om, tc.sampling_intvl


def _power_rows(om, path):
    lines = [l.split() for l in open(path) if l.strip()]
    names, rows = lines[0], [list(map(float, l)) for l in lines[1:]]
    col = {}
    base = 0
    for lay in om.layers:
        if lay.has_power:
            for k, u in enumerate(lay.units):
                col[u.name] = base + k
        base += len(lay.units)
    P = np.zeros((len(rows), om.n_vector()))
    for r, vals in enumerate(rows):
        for nm, v in zip(names, vals):
            P[r, col[nm]] = v
    return names, P


def test_example5_microchannel_39_through_cli(gpu, tmp_path):
    d = str(tmp_path / "ex5")
    shutil.copytree(EX5, d)
    os.makedirs(os.path.join(d, "outputs"))
    common = ["-c", "example.config", "-p", "example.ptrace", "-materials_file", "example.materials",
              "-grid_layer_file", "example.lcf", "-model_type", "grid", "-detailed_3D", "on",
              "-use_microchannels", "1"]
    _run_cli(common + ["-grid_steady_file", "outputs/example.grid.steady", "-steady_file",
                       "outputs/example.steady"], d)
    shutil.copy(os.path.join(d, "outputs", "example.steady"), os.path.join(d, "example.init"))
    _run_cli(common + ["-init_file", "example.init", "-o", "outputs/example.transient",
                       "-transient_solver", "be"], d)
    om, dt = _example5_oracle(d)
    assert (om.nr, om.nc) == (39, 39) and not om.symmetric
    names, P = _power_rows(om, os.path.join(d, "example.ptrace"))
    # steady: the average power of the trace (hotspot.c:611-622)
    To, Tg = om.steady(P.mean(axis=0))
    got = _read_block_file(os.path.join(d, "outputs", "example.steady"))
    idx = _entry_index(om)
    worst = max(abs(v - To[idx[k]]) for k, v in got.items() if k in idx)
    assert set(got) == set(idx)
    assert worst <= 0.005 + 1e-4, worst    # the file prints %.2f
    # grid steady file: layer by layer, "cell value"
    grid_lines = [l.split() for l in open(os.path.join(d, "outputs", "example.grid.steady"))
                  if l.strip() and not l.startswith("Layer")]
    vals = np.array([float(t[1]) for t in grid_lines])
    assert vals.size == om.ncell
    assert np.abs(vals - Tg[:om.ncell]).max() <= 0.005 + 1e-4
    # transient: one backward-Euler step per row from the printed steady state
    init = np.zeros(om.n_vector())
    for k, v in _read_block_file(os.path.join(d, "example.init")).items():
        if k in idx:
            init[idx[k]] = v
    y = om.b2g(init, power=False)
    rows = [l.split() for l in open(os.path.join(d, "outputs", "example.transient")) if l.strip()]
    assert rows[0] == names
    for r, vals in enumerate(rows[1:]):
        y = om.advance_be(y, om.power_grid(P[r]), dt)
        Tb = om.g2b(y)
        got_r = np.array(list(map(float, vals)))
        want_r = np.array([Tb[_col(om, nm)] for nm in names])
        assert np.abs(got_r - want_r).max() <= 0.005 + 1e-4, r
    assert len(rows) - 1 == P.shape[0]


def _entry_index(om):
    """Names of a layer-file model's temperature file entries (the
    reference's dump_temp for the grid model): layer_<n>_<unit> for the chip
    layers, hsp_<unit> and hsink_<unit> for the spreader and the sink, and
    inode_<k> for the package nodes."""
    idx, base = {}, 0
    for n, lay in enumerate(om.layers):
        pre = "hsp_" if n == om.spidx else "hsink_" if n == om.hsidx else f"layer_{n}_"
        for k, u in enumerate(lay.units):
            idx[pre + u.name] = base + k
        base += len(lay.units)
    for k in range(om.nx):
        idx[f"inode_{k}"] = base + k
    return idx


def _col(om, name):
    base = 0
    for lay in om.layers:
        if lay.has_power:
            for k, u in enumerate(lay.units):
                if u.name == name:
                    return base + k
        base += len(lay.units)
    raise KeyError(name)
