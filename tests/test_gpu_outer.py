"""The outer loops around the steady solve of the grid model, on the GPU.

The temperature-leakage fixed point (steady_state_temp, temperature.c:
708-746): each steady solve's power adds the leakage of every power-layer
block at the last solve's block temperatures (calc_leakage,
temperature.c:910-930), until no block rises by 0.01 K.
  * Through the command line on the reference's example 2 (ev6, 64 x 64):
    the block and grid steady files against the reference command line's own
    outputs on the same inputs, within one printed digit.  The reference's
    steady sweeps stop about 1e-3 K early (1e-6 K per sweep), and the fixed
    point moves with them, so fewer prints agree exactly than without
    leakage.
  * Against the oracle: the same loop around the oracle's converged steady
    solve (restated here from temperature.c:708-746), within 1e-4 K.
  * In the batched engine, beside jobs without leakage: each result equals
    steady_state_temp's bit for bit (the engine warm-starts every re-solve
    from the last, as the single-model loop does), with the same number of
    solves; a fixed point that runs away fails its job with the reference's
    message and leaves the other jobs alone.
"""
import os
import shutil

import numpy as np
import pytest

from hotspot_amd import cases, engine
from helpers import check_error, make, oracle_for
from test_gpu_cli import CLI_DIR, compare, read_text, run_cli

pytestmark = pytest.mark.gpu


def test_cli_grid_leakage_matches_reference(gpu, tmp_path):
    src = os.path.join(CLI_DIR, "ev6")
    for f in os.listdir(src):
        if os.path.isfile(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), tmp_path / f)
    run_cli("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type grid "
            "-leakage_used 1 -steady_file leak.steady -grid_steady_file leak.grid.steady", tmp_path)
    for f in ("leak.steady", "leak.grid.steady"):
        m, ex = compare(open(tmp_path / f).read(), read_text(os.path.join(src, "expected", f)), 0.3)
        print(f, "max |diff|", m, "exact", ex)


def leakage_power(om, P, T):
    """power + calc_leakage of every power-layer block at T (oracle units)."""
    out = np.zeros_like(P)
    base = 0
    for lay in om.layers:
        n = len(lay.units)
        if lay.has_power:
            for k, u in enumerate(lay.units):
                out[base + k] = P[base + k] + 1.5e4 * u.height * u.width * np.exp(0.036 * (T[base + k] - 383.15))
        base += n
    return out


def oracle_leakage_fixed_point(om, P):
    T = np.zeros(om.n_vector())
    power_blocks = np.zeros(om.n_vector(), dtype=bool)
    base = 0
    for lay in om.layers:
        if lay.has_power:
            power_blocks[base:base + len(lay.units)] = True
        base += len(lay.units)
    for it in range(101):
        Pn = leakage_power(om, P, T)
        Tn, _ = om.steady(Pn)
        rise = float(np.max((Tn - T)[power_blocks], initial=0.0))
        T = Tn
        if rise < 0.01:
            return T, it + 1
    raise AssertionError("no convergence")


def _leak_cases():
    c1 = cases.default_stack_case(64, leakage_used=1)
    p1 = [cases.rng(21).uniform(1.0, 6.0, 16)]
    c2, p2 = cases.stacked_case(64, seed=22, n_dies=2, blocks_per_die=16)
    c2.config["leakage_used"] = 1
    return [("default-64", c1, p1), ("stack-64", c2, p2)]


def _steady_single(lib, run, P, prec):
    from helpers import set_options
    set_options(lib, run, precision=prec, steady_solver=1 if prec == 1 else 0)
    tp, ta = run.vec(np.zeros(run.n))
    pp, pa = run.vec(P)
    lib.lib.steady_state_temp(run.model, pp, tp)
    check_error(lib)
    return ta.copy()


@pytest.mark.parametrize("name,case,pw", _leak_cases(), ids=[c[0] for c in _leak_cases()])
def test_grid_leakage_matches_oracle(prod, gpu, name, case, pw):
    run = make(prod, case)
    P = run.power_vector(pw)
    om = oracle_for(run)
    want, iters = oracle_leakage_fixed_point(om, P)
    assert iters >= 2
    got = _steady_single(prod, run, P, 0)
    assert np.abs(got - want).max() <= 1e-4
    run.free()


@pytest.mark.parametrize("form", [(engine.FP64, engine.RK4), (engine.MIXED, engine.TAYLOR)],
                         ids=["fp64", "mixed"])
def test_engine_leakage_equals_single_model(prod, gpu, form):
    specs = _leak_cases()
    runs = [make(prod, c) for _, c, _ in specs]
    plain = make(prod, cases.default_stack_case(32))
    Ps = [r.power_vector(pw) for r, (_, _, pw) in zip(runs, specs)]
    Pp = plain.power_vector([cases.rng(23).uniform(0.5, 5.0, 16)])
    eng = engine.Engine(prod, *form)
    ids = [eng.submit(r.model, P) for r, P in zip(runs, Ps)]
    ip = eng.submit(plain.model, Pp)
    # a leakage job with a transient phase after its fixed point
    seq = np.vstack([Ps[0][None, :], np.array(cases.jittered_powers(Ps[0], 5, seed=3))])
    it = eng.submit(runs[0].model, seq, n_steps=5, dt=3.333e-6)
    eng.run()
    served = [eng.poll(i) for i in ids]
    infos = [eng.info(i) for i in ids]
    served_plain = eng.poll(ip)
    served_seq = eng.result_seq(it)
    info_t = eng.info(it)
    eng.close()
    prec = 1 if form[0] == engine.MIXED else 0
    for r, P, got, info in zip(runs, Ps, served, infos):
        want = _steady_single(prod, r, P, prec)
        assert np.array_equal(got, want)
        assert info.steady_solves >= 2
    assert eng_plain_equal(prod, plain, Pp, prec, served_plain)
    assert info_t.steady_solves == infos[0].steady_solves
    # the transient phase starts from the fixed point's block temperatures
    e1 = engine.Engine(prod, *form)
    j = e1.submit(runs[0].model, seq[1:], init=served[0], n_steps=5, dt=3.333e-6)
    e1.run()
    assert np.array_equal(e1.result_seq(j), served_seq)
    e1.close()
    for r in runs + [plain]:
        r.free()


def eng_plain_equal(lib, run, P, prec, got):
    return np.array_equal(_steady_single(lib, run, P, prec), got)


def test_engine_leakage_runaway_fails_its_job_only(prod, gpu):
    """Leakage that grows faster than the sink removes heat (80 W through a
    5 K/W sink: about 720 K after the first solve, where each block leaks
    tens of kW): the second solve rises by more than 500 K and the job fails
    with the reference's message; a job beside it finishes."""
    hot = make(prod, cases.default_stack_case(32, leakage_used=1, r_convec=5.0))
    ok = make(prod, cases.default_stack_case(32))
    Ph = hot.power_vector([np.full(16, 5.0)])
    Po = ok.power_vector([cases.rng(24).uniform(0.5, 5.0, 16)])
    eng = engine.Engine(prod, engine.FP64, engine.RK4)
    jh = eng.submit(hot.model, Ph)
    jo = eng.submit(ok.model, Po)
    eng.run()
    assert eng.info(jh).state == engine.FAILED
    with pytest.raises(RuntimeError, match="thermal runaway"):
        eng.poll(jh)
    assert eng.info(jo).state == engine.DONE and eng.poll(jo) is not None
    eng.close()
    hot.free()
    ok.free()
