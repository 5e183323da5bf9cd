"""The oracle against the reference's golden vectors, and against what the
mathematics fixes (symmetry, zero row sums, energy balance, linearity, the
maximum principle, block-map occupancies, the exact ODE solution).

These are the checks that pin each oracle function; a restatement that drops
a term, flips a sign or swaps an index fails at least one of them.
"""
import os
import tempfile

import numpy as np
import pytest
import scipy.sparse as sp

import hs_oracle as O
from hotspot_amd import cases, iface

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden_cases():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_goldens", os.path.join(GOLDEN, "make_goldens.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


MG = _golden_cases()
NAMES = [c[0] for c in MG.golden_cases()] + ["microchannel16"]


def _load(name):
    """(oracle model, block power, fixture) for a golden case."""
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    tmp = tempfile.mkdtemp()
    if name == "microchannel16":
        _, case, pw = MG.microchannel_case(tmp)
    else:
        case, pw = {c[0]: (c[1], c[2]) for c in MG.golden_cases()}[name]
    m = case.materialize(tmp)
    om = O.OracleModel(m["config"], flp=m["flp"], detailed_3D=case.detailed_3D,
                       use_microchannels=case.use_microchannels, microchannel=case.microchannel)
    P = np.zeros(om.n_vector())
    base, k = 0, 0
    for lay in om.layers:
        if lay.has_power:
            P[base:base + len(lay.units)] = pw[k]
            k += 1
        base += len(lay.units)
    return om, P, d


@pytest.fixture(scope="module", params=NAMES)
def golden(request):
    return (request.param,) + _load(request.param)


def test_inputs_unchanged(golden):
    name, om, P, d = golden
    geo = np.array([[u.width, u.height, u.leftx, u.bottomy] for lay in om.layers for u in lay.units])
    assert np.array_equal(geo, d["inputs_geometry"])
    assert np.array_equal(P, d["inputs_power"])


def test_layers_and_package(golden):
    name, om, P, d = golden
    for f in ("k", "thickness", "sp", "rx", "ry", "rz", "c"):
        mine = np.array([getattr(l, f) for l in om.layers])
        assert np.allclose(mine, d["layer_" + f], rtol=1e-14, atol=0), f
    mine = np.array([om.pack[f] for f in iface.PACKAGE_FIELDS])
    assert np.allclose(mine, d["pack"], rtol=1e-14, atol=0)


def test_block_grid_maps(golden):
    name, om, P, d = golden
    rows = []
    for n, lay in enumerate(om.layers):
        g2b = om.maps[lay.map_id]["g2b"]
        for u in range(len(lay.units)):
            rows.append([n, u] + list(g2b[u]))
    assert np.array_equal(np.array(rows), d["g2b"])
    Pg = om.power_grid(P)
    assert np.array_equal(Pg[:om.ncell], d["b2g_power"][:om.ncell])


def test_operator_against_reference_slope(golden):
    """slope_fn_grid of the reference on seeded temperatures pins G, b, C."""
    name, om, P, d = golden
    # the fixture's slope was taken with the case's own power on the grid
    s = om.slope(d["slope_T"], d["b2g_power"] * np.r_[np.ones(om.ncell), np.zeros(om.nx)])
    assert np.abs(s - d["slope"]).max() <= 1e-11 * np.abs(d["slope"]).max()


def test_steady_against_converged_reference(golden):
    name, om, P, d = golden
    if "steady_conv" not in d:
        pytest.skip("the reference has no steady solve for microchannel layers without SuperLU")
    Tb, Tg = om.steady(P)
    assert np.abs(Tg - d["steady_conv"]).max() <= 1e-8
    # the stock output stops sweeping at a 1e-6 K change per sweep
    assert np.abs(Tb - d["steady_stock"]).max() <= 5e-3


def test_transient_against_reference(golden):
    name, om, P, d = golden
    Tg = om.b2g(d["trans_T0"], power=False)
    for k, p in enumerate(d["trans_P"]):
        Tg, _ = om.advance(Tg, om.power_grid(p), 3.333e-6)
        assert np.abs(om.g2b(Tg) - d["trans"][k]).max() <= 1e-8


# ---- what the mathematics fixes --------------------------------------------
def _default(n=16, **kw):
    case = cases.default_stack_case(n, **kw)
    m = case.materialize(tempfile.mkdtemp())
    return O.OracleModel(m["config"], flp=m["flp"])


def test_operator_symmetric_m_matrix():
    om = _default()
    G = om.G
    assert abs(G - G.T).max() <= 1e-12 * abs(G).max()
    off = G - sp.diags(G.diagonal())
    assert off.max() <= 0.0                       # non-positive off-diagonals
    ground = np.asarray(G.sum(axis=1)).ravel()    # row sums = ambient conductance
    assert (ground >= -1e-12 * abs(G).max()).all()
    assert np.allclose(ground, om.b / om.cfg["ambient"], rtol=1e-9, atol=1e-12)


def test_energy_balance_and_maximum_principle():
    om = _default()
    P = np.zeros(om.n_vector())
    P[:16] = cases.rng(9).uniform(0.5, 5, 16)
    _, Tg = om.steady(P)
    amb = om.cfg["ambient"]
    # total power leaves through the ambient conductances
    lost = np.sum(om.b / amb * (Tg - amb))
    assert abs(lost - P[:16].sum()) <= 1e-9 * P[:16].sum()
    assert Tg.min() >= amb - 1e-9
    # zero power: ambient everywhere
    _, T0 = om.steady(np.zeros(om.n_vector()))
    assert np.abs(T0 - amb).max() <= 1e-9


def test_linearity():
    om = _default()
    g = cases.rng(4)
    P1 = np.zeros(om.n_vector()); P1[:16] = g.uniform(0, 3, 16)
    P2 = np.zeros(om.n_vector()); P2[:16] = g.uniform(0, 3, 16)
    amb = om.cfg["ambient"]
    _, T1 = om.steady(P1)
    _, T2 = om.steady(P2)
    _, T3 = om.steady(2 * P1 + 3 * P2)
    assert np.abs((T3 - amb) - (2 * (T1 - amb) + 3 * (T2 - amb))).max() <= 1e-9


def test_occupancies_sum_to_one():
    """The reference's own (unused) self-check test_b2gmap
    (temperature_grid.c:3752-3773): occupancies of every cell sum to 1."""
    case = cases.default_stack_case(16, nbx=3, nby=5)
    m = case.materialize(tempfile.mkdtemp())
    om = O.OracleModel(m["config"], flp=m["flp"])
    mp = om.maps[0]
    s = np.zeros(16 * 16)
    np.add.at(s, mp["cell"], mp["occ"])
    assert np.allclose(s, 1.0, atol=1e-12)
    # block-uniform temperatures come back unchanged through b2g and g2b
    Tb = np.zeros(om.n_vector())
    Tb[:15] = cases.rng(2).uniform(300, 400, 15)
    for n in range(1, om.nl):
        Tb[15 * n:15 * (n + 1)] = Tb[:15]
    back = om.g2b(om.b2g(Tb, power=False))
    # center mapping of a block whose centre cells are fully inside it
    inner = [u for u in range(15) if (mp["g2b"][u][1] - mp["g2b"][u][0]) >= 3 and (mp["g2b"][u][3] - mp["g2b"][u][2]) >= 3]
    assert np.allclose(back[inner], Tb[inner], atol=1e-9)


def test_power_conserved_by_mapping():
    case, pw = cases.stacked_case(32, seed=2, blocks_per_die=16)
    m = case.materialize(tempfile.mkdtemp())
    om = O.OracleModel(m["config"], flp=m["flp"])
    P = np.zeros(om.n_vector())
    base, k = 0, 0
    for lay in om.layers:
        if lay.has_power:
            P[base:base + len(lay.units)] = pw[k]
            k += 1
        base += len(lay.units)
    assert abs(om.power_grid(P).sum() - P.sum()) <= 1e-10 * P.sum()


def test_exact_transient_against_fine_rk4_and_decay():
    om = _default()
    P = np.zeros(om.n_vector()); P[:16] = 2.0
    Pg = om.power_grid(P)
    y0 = np.full(om.N, om.cfg["ambient"] + 5.0)
    ye = om.advance_exact(y0, Pg, 3.333e-6)
    yf = y0.copy()
    h = 3.333e-6 / 200
    for _ in range(200):
        yf = om.rk4_core(yf, om.slope(yf, Pg), Pg, h)
    assert np.abs(ye - yf).max() <= 1e-9
    # with no power a uniform excess decays monotonically and stays above ambient
    amb = om.cfg["ambient"]
    yz = om.advance_exact(y0, np.zeros(om.N), 1e-3)
    assert (yz <= y0 + 1e-12).all() and (yz >= amb - 1e-9).all()
    # steady state is a fixed point of the transient
    Ts = om.steady_grid(Pg)
    assert np.abs(om.advance_exact(Ts, Pg, 1e-4) - Ts).max() <= 1e-8


def test_rk4_fixed_point_and_order():
    om = _default()
    amb = om.cfg["ambient"]
    y = np.full(om.N, amb)
    y2, calls = om.advance(y, np.zeros(om.N), 3.333e-6)
    assert np.abs(y2 - amb).max() <= 1e-12 and calls >= 1
    # fourth order: halving the step cuts one RK4 step's error ~16x
    P = np.zeros(om.n_vector()); P[:16] = 3.0
    Pg = om.power_grid(P)
    y0 = np.full(om.N, amb + 3.0)
    ex = om.advance_exact(y0, Pg, 2e-5)
    e1 = np.abs(om.rk4_core(y0, om.slope(y0, Pg), Pg, 2e-5) - ex).max()
    yh = om.rk4_core(y0, om.slope(y0, Pg), Pg, 1e-5)
    e2 = np.abs(om.rk4_core(yh, om.slope(yh, Pg), Pg, 1e-5) - ex).max()
    assert 8 < e1 / e2 < 40


def test_microchannel_direct_solve_satisfies_the_operator():
    """The reference cannot solve microchannel steady states without SuperLU;
    the oracle's direct solve of its (golden-pinned) non-symmetric operator is
    pinned by its residual and by energy balance with the coolant."""
    om, P, d = _load("microchannel16")
    assert not om.symmetric
    Pg = om.power_grid(P)
    Tg = om.steady_grid(Pg)
    r = om.residual(Tg, Pg)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(Pg + om.b)
    asym = abs(om.G - om.G.T).max()
    assert asym > 1e-6 * abs(om.G).max()


def test_backward_euler_limits():
    """advance_be: its fixed point is the steady state (h -> infinity), and its error against the exact solution falls linearly
    with the step (first order: halving h halves the error); a step without
    the 1/h on C T would break both."""
    om = _default()
    amb = om.cfg["ambient"]
    P = np.zeros(om.n_vector()); P[:16] = 3.0
    Pg = om.power_grid(P)
    y0 = np.full(om.N, amb + 3.0)
    tinf = spsolve_t(om, Pg)
    assert np.abs(om.advance_be(y0, Pg, 1e9) - tinf).max() <= 1e-6
    assert np.abs(om.advance_be(tinf, Pg, 1e-5) - tinf).max() <= 1e-9
    dt = 2e-5
    ex = om.advance_exact(y0, Pg, dt)
    errs = []
    for n in (4, 8, 16):
        y = y0.copy()
        for _ in range(n):
            y = om.advance_be(y, Pg, dt / n)
        errs.append(np.abs(y - ex).max())
    assert 1.7 < errs[0] / errs[1] < 2.3 and 1.7 < errs[1] / errs[2] < 2.3


def spsolve_t(om, Pg):
    import scipy.sparse.linalg as spla
    return spla.spsolve(om.G.tocsc(), Pg + om.b)


def _be_with_caps(om, Tg, Pg, h, cbe, G):
    import scipy.sparse.linalg as spla
    A = (sp.diags(cbe / h) + G).tocsc()
    return spla.spsolve(A, cbe / h * Tg + Pg + om.b)


def test_backward_euler_secondary_model():
    """advance_be on the secondary-path model (the SuperLU build's matrix,
    temperature_grid.c:6354, :6440 and build_diagonal_matrix :6750-6768):
    the secondary package nodes carry no capacitance there, so after every
    step they satisfy their steady equations exactly, whatever the state
    before it; the other rows satisfy C (T' - T)/h = P + b - G T'; and the
    step's fixed point for h -> infinity is the steady state of the steady
    operator G (with the solder-substrate links), not of the transient Gt.
    The same checks fail on two planted mistakes: secondary nodes with their
    slope_fn_pack capacitances, and the transient operator Gt in place of G."""
    om, P, _ = _load("secondary16")
    assert om.secondary
    Pg = om.power_grid(P)
    rhs = Pg + om.b
    sec = np.arange(om.ncell + O.EXTRA, om.N)
    prim = np.arange(om.ncell + O.EXTRA)
    y = om.b2g(np.full(om.n_vector(), om.cfg["ambient"] + 20.0), power=False)
    y[sec] += np.linspace(-5.0, 5.0, len(sec))      # secondary nodes off their balance
    h = 3.333e-6

    def violations(ynew, yold, G, cbe):
        r = rhs - G @ ynew
        alg = np.abs(r[sec]).max() / np.abs(rhs[sec] + (G @ ynew)[sec]).max()
        dyn = (np.abs(cbe[prim] * (ynew[prim] - yold[prim]) / h - r[prim]).max()
               / np.abs(cbe[prim] * ynew[prim] / h).max())
        return alg, dyn

    for k in range(3):
        y1 = om.advance_be(y, Pg, h)
        alg, dyn = violations(y1, y, om.G, np.where(np.arange(om.N) < om.ncell + O.EXTRA, om.Cvec, 0.0))
        assert alg <= 1e-9 and dyn <= 1e-9, (k, alg, dyn)
        y = y1
    tinf = spsolve_t(om, Pg)
    assert np.abs(om.advance_be(y, Pg, 1e9) - tinf).max() <= 1e-6
    # planted mistake 1: the secondary nodes keep their transient capacitances
    y0 = om.b2g(np.full(om.n_vector(), om.cfg["ambient"] + 20.0), power=False)
    y0[sec] += np.linspace(-5.0, 5.0, len(sec))
    bad = _be_with_caps(om, y0, Pg, h, om.Cvec, om.G)
    alg_bad, _ = violations(bad, y0, om.G, om.Cvec)
    assert alg_bad > 1e-3
    # planted mistake 2: the transient operator in the implicit step
    cbe = np.where(np.arange(om.N) < om.ncell + O.EXTRA, om.Cvec, 0.0)
    bad2 = _be_with_caps(om, y0, Pg, 1e9, cbe, om.Gt)
    assert np.abs(bad2 - tinf).max() > 1e-3


def _mc_oracle(width_cells, superlu):
    d = tempfile.mkdtemp()
    case, _ = cases.microchannel_case(16, 5, d, width_cells=width_cells, pitch_cells=width_cells + 2)
    m = case.materialize(d)
    return O.OracleModel(m["config"], lcf=m["lcf"], detailed_3D=True, use_microchannels=True,
                         microchannel=case.microchannel, superlu_outlets=superlu), d


def test_superlu_outlet_rule():
    """The SuperLU build's outlet rows (temperature_grid.c:5313-5345): an
    outlet takes its advection from its first fluid neighbour alone, in the
    order N, S, W, E.  Where every outlet has one fluid neighbour (one-cell
    channels between solid rows, as examples/example5) the operator equals
    the slope's exactly.  Where outlets touch (two-cell channels), an outlet
    whose first fluid neighbour is another outlet (no flow between them)
    loses its advection: the difference of the two operators is, in that row
    alone, the slope's terms restated here -- -Cv Q/2 to the upstream cell,
    Cv Q/2 on the diagonal.  A rule that took the upstream (W) neighbour
    first would equal the slope's there and leave that row difference zero."""
    a, _ = _mc_oracle(1, False)
    b, _ = _mc_oracle(1, True)
    assert abs(a.G - b.G).max() == 0.0
    assert np.array_equal(a.b, b.b)
    s, _ = _mc_oracle(2, False)
    r, _ = _mc_oracle(2, True)
    D = (s.G - r.G).tocsr()
    lay = next(n for n in range(s.nl) if s.layers[n].is_microchannel)
    mc = s.layers[lay].mc
    types, p, hC, Cv = mc["cell_types"], mc["pressure"], mc["hydroC"], mc["coolant_capac"]
    nr, nc = types.shape
    rows = set(D.nonzero()[0])
    outlets = 0
    for i in range(nr):
        for j in range(nc):
            me = (lay * nr + i) * nc + j
            if types[i, j] != O.OUTLET:
                assert me not in rows
                continue
            outlets += 1
            # the two-cell channels' outlets: the first fluid neighbour (N or
            # S) is the other outlet of the channel, the upstream cell is W
            assert types[i - 1, j] == O.OUTLET or types[i + 1, j] == O.OUTLET
            q = (p[i, j - 1] - p[i, j]) * hC
            assert q > 0
            row = D.getrow(me).toarray().ravel()
            want = np.zeros_like(row)
            want[me - 1] = -Cv * q / 2.0
            want[me] = -Cv * q / 2.0 + Cv * q     # the edge outlet's outflow
            assert np.allclose(row, want, rtol=1e-12, atol=1e-15 * abs(s.G).max())
    assert outlets >= 4
