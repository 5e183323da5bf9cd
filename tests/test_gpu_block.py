"""The block model on the GPU (hs_block.hip) against the reference library
compiled from its sources (oracle/_ref) and the oracle, on the same inputs.

* LU factors: the device factorisation repeats the reference's partial-
  pivoting elimination operation by operation, so the factors and the
  permutation are the reference's to the bit.
* Steady: forward and backward substitution; within 1e-9 K of the
  reference (only the order of the backward sums differs).
* Transient: compute_temp's adaptive RK4 over 3.333 us sampling intervals
  and over one 1 s interval (about 11 000 steps): bit-identical to the
  reference.  The attempts repeat the reference's arithmetic and sums in
  its order, and the step-size controller uses the C library's pow as the
  reference does; over long intervals anything less drifts onto another
  step sequence (the oracle, whose numpy sums round differently, ends
  4e-3 K away after 1 s, as far as the reference is from the exact
  solution).
* The leakage fixed point and the singular-matrix error of the reference's
  LU.
"""
import ctypes as C

import numpy as np
import pytest

import hs_oracle as O
from block_helpers import EV6, BlockRun, synthetic_flp
from helpers import HsStats, bind_extensions

pytestmark = pytest.mark.gpu

CASES = {"ev6": (None, {}), "ev6_omit_lateral": (None, {"block_omit_lateral": 1}),
         "slice12": ((12, 3), {}), "slice64": ((64, 9), {"r_convec": 0.3})}


def case_flp(name, tmp_path):
    spec = CASES[name][0]
    return EV6 if spec is None else synthetic_flp(str(tmp_path / f"{name}.flp"), spec[0], seed=spec[1])


def seeded_power(n, seed):
    return np.random.default_rng(seed).uniform(0.5, 5.0, n)


@pytest.mark.parametrize("case", list(CASES))
def test_block_factors_and_steady_match_reference(gpu, prod, ref, case, tmp_path):
    path, cfg = case_flp(case, tmp_path), CASES[case][1]
    print("block steady:", case, flush=True)
    a = BlockRun(prod, path, **cfg)
    r = BlockRun(ref, path, **cfg)
    try:
        # populate_R factorised B on the device
        assert np.array_equal(a.matrix("lu"), r.matrix("lu"))
        assert np.array_equal(np.ctypeslib.as_array(a.b.p, shape=(a.N,)), np.ctypeslib.as_array(r.b.p, shape=(r.N,)))
        pw = seeded_power(a.n_units, 21)
        got, want = a.steady(pw), r.steady(pw)
        d = np.abs(got - want).max()
        print("  max |steady - reference|", d)
        assert d <= 1e-9
        o = O.OracleBlockModel(dict(cfg), flp=path)
        assert np.abs(got - o.steady(pw)).max() <= 1e-9
    finally:
        a.close()
        r.close()


@pytest.mark.parametrize("case", ["ev6", "slice64"])
def test_block_transient_matches_reference(gpu, prod, ref, case, tmp_path):
    path, cfg = case_flp(case, tmp_path), CASES[case][1]
    print("block transient:", case, flush=True)
    lib = bind_extensions(prod)
    a = BlockRun(prod, path, **cfg)
    r = BlockRun(ref, path, **cfg)
    try:
        g = np.random.default_rng(13)
        base = seeded_power(a.n_units, 2)
        rows = [base * g.uniform(0.8, 1.2, a.n_units) * 4 for _ in range(40)]
        init = r.steady(base)
        got = a.transient(rows, init, 3.333e-6)
        want = r.transient(rows, init, 3.333e-6)
        d = np.abs(got - want).max()
        print("  max |transient - reference| over 40 intervals", d)
        assert np.array_equal(got, want)
        # RK4 steps of the last interval, against the oracle's count
        o = O.OracleBlockModel(dict(cfg), flp=path)
        _, calls = o.transient(rows[-1], want[-2], 3.333e-6)
        st = HsStats()
        lib.lib.hs_get_stats(a.model, C.byref(st))
        assert st.rk4_calls == calls
    finally:
        a.close()
        r.close()


def test_block_long_interval_matches_reference(gpu, prod, ref):
    """One compute_temp of 1 s from ambient: ~11 000 RK4 steps.  Over
    hundreds of steps the adaptive step sequence depends on the last bits of
    every step-doubling error estimate: the result is bit-identical to the
    reference's, and both lie within the controller's 0.01 K per step of the
    exact solution T_inf + exp(-A^-1 B t)(T0 - T_inf)."""
    from scipy.linalg import expm
    print("block transient: ev6, one 1 s interval", flush=True)
    lib = bind_extensions(prod)
    a = BlockRun(prod, EV6)
    r = BlockRun(ref, EV6)
    try:
        pw = seeded_power(a.n_units, 8)
        init = np.full(a.N, a.cfg.ambient)
        got = a.transient([pw], init, 1.0)[0]
        want = r.transient([pw], init, 1.0)[0]
        st = HsStats()
        lib.lib.hs_get_stats(a.model, C.byref(st))
        o = O.OracleBlockModel(flp=EV6)
        T = o.steady(pw)
        exact = T + expm(-1.0 * o.Cm) @ (init - T)
        d, de, dr = np.abs(got - want).max(), np.abs(got - exact).max(), np.abs(want - exact).max()
        print("  RK4 steps", st.rk4_calls, "|gpu - reference|", d, "|gpu - exact|", de, "|reference - exact|", dr)
        assert st.rk4_calls > 10000
        assert np.array_equal(got, want) and de <= 1e-2
    finally:
        a.close()
        r.close()


def test_block_leakage_steady_matches_reference(gpu, prod, ref):
    print("block steady with leakage: ev6", flush=True)
    a = BlockRun(prod, EV6, leakage_used=1)
    r = BlockRun(ref, EV6, leakage_used=1)
    try:
        pw = seeded_power(a.n_units, 4)
        # the fixed point starts from the temperature vector's contents
        outs = []
        for run in (a, r):
            pp, p = run.vec(pw)
            tp, t = run.vec(np.full(run.N, 330.0))
            run.lib.lib.steady_state_temp(run.model, pp, tp)
            outs.append(t.copy())
        d = np.abs(outs[0] - outs[1]).max()
        print("  max |diff|", d)
        assert d <= 1e-8
    finally:
        a.close()
        r.close()


def test_block_singular_matrix_error(gpu, prod):
    """An undefined convection resistance (natural convection of a sink at
    ambient gives 0/0) leaves columns of B without a usable pivot: the LU
    stops in populate_R with the reference's message (whose library exits
    there, so only the message is compared)."""
    lib = bind_extensions(prod)
    buf = C.create_string_buffer(512)
    lib.lib.hs_last_error(buf, 512)
    a = BlockRun(prod, EV6, r_convec=float("nan"))
    try:
        assert lib.lib.hs_last_error(buf, 512) == 1
        assert "singular matrix in lupdcmp" in buf.value.decode()
    finally:
        a.close()
