"""The product library's host-side model (configuration, floorplan and layer
parsing, layer stack, block-grid maps, layer and package R/C) against the
golden vectors of the reference library, read through the reference's own
struct layouts.  No GPU needed: the device is first touched by a solve."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from hotspot_amd import cases, iface
from test_oracle import MG, NAMES, GOLDEN


@pytest.fixture(scope="module")
def lib(prod):
    return prod


def _case(name, tmp):
    if name == "microchannel16":
        _, case, pw = MG.microchannel_case(tmp)
        return case, pw
    return {c[0]: (c[1], c[2]) for c in MG.golden_cases()}[name]


@pytest.mark.parametrize("name", NAMES)
def test_host_model_matches_reference(lib, name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    tmp = tempfile.mkdtemp()
    case, pw = _case(name, tmp)
    run = cases.Run(lib, case, tmp)
    g = run.grid
    cfg = d["config"]
    assert [g.n_layers, g.rows, g.cols, g.config.model_secondary, g.map_mode, g.has_lcf,
            g.total_n_blocks] == list(cfg)
    for f in ("k", "thickness", "sp", "rx", "ry", "rz", "c", "has_lateral", "has_power"):
        mine = np.array([getattr(g.layers[n], f) for n in range(g.n_layers)], dtype=float)
        assert np.array_equal(mine, d["layer_" + f]), f
    pk = np.array([getattr(g.pack, f) for f in iface.PACKAGE_FIELDS])
    assert np.array_equal(pk, d["pack"])
    rows = []
    for n in range(g.n_layers):
        for u in range(g.layers[n].flp.contents.n_units):
            r = g.layers[n].g2bmap[u]
            rows.append([n, u, r.i1, r.i2, r.j1, r.j2])
    assert np.array_equal(np.array(rows), d["g2b"])
    assert np.array_equal(run.power_vector(pw), d["inputs_power"])
    run.free()


def test_vectors_and_text_formats(lib):
    case = cases.default_stack_case(16)
    tmp = tempfile.mkdtemp()
    run = cases.Run(lib, case, tmp)
    L = lib.lib
    assert run.n == 4 * 16 + 12
    p, a = run.vec()
    L.set_temp(run.model, p, 333.0)
    assert np.all(a == 333.0)
    a[:] = np.arange(run.n) + 300.0
    f = os.path.join(tmp, "t.txt")
    L.dump_temp(run.model, p, f.encode())
    lines = open(f).read().split("\n")
    assert lines[0].startswith("b0_0\t300.00") and lines[16].startswith("iface_b0_0\t316.00")
    assert lines[run.n - 12].startswith("inode_0\t")
    q, b = run.vec()
    L.read_temp(run.model, q, f.encode(), 0)
    assert np.allclose(b, np.round(a, 2))
    assert abs(L.find_max_temp(run.model, p) - 315.0) < 1e-12
    assert abs(L.find_avg_temp(run.model, p) - np.mean(a[:16])) < 1e-12
    run.free()


def _expect_error(lib, fn):
    buf = C.create_string_buffer(512)
    lib.lib.hs_last_error(buf, 512)
    out = fn()
    assert lib.lib.hs_last_error(buf, 512) == 1
    return out, buf.value.decode()


def test_errors_are_reported_not_fatal(lib):
    L = lib.lib
    L.hs_last_error.argtypes = [C.c_char_p, C.c_int]
    L.hs_last_error.restype = C.c_int
    cfg = lib.default_config(model_type="grid", grid_rows=0, grid_cols=64)
    tmp = tempfile.mkdtemp()
    fp = os.path.join(tmp, "a.flp")
    cases.write_flp(fp, cases.tiled_floorplan(2, 2, 0.01, 0.01))
    flp = L.read_flp(fp.encode(), 0, 0)
    m, msg = _expect_error(lib, lambda: L.alloc_RC_model(C.byref(cfg), flp, None, None, 0, 0))
    assert not m and "must be positive" in msg
    # any positive size is taken (the reference's SuperLU build takes them)
    cfg = lib.default_config(model_type="grid", grid_rows=39, grid_cols=48)
    m = L.alloc_RC_model(C.byref(cfg), flp, None, None, 0, 0)
    assert m
    L.populate_R_model(m, flp)
    assert m.contents.grid.contents.rows == 39 and m.contents.grid.contents.cols == 48
    L.delete_RC_model(m)
    _, msg = _expect_error(lib, lambda: L.read_flp(os.path.join(tmp, "missing.flp").encode(), 0, 0))
    assert "missing.flp" in msg
    big = os.path.join(tmp, "big.flp")
    cases.write_flp(big, cases.tiled_floorplan(2, 2, 0.05, 0.05))
    flp2 = L.read_flp(big.encode(), 0, 0)
    cfg = lib.default_config(model_type="grid", grid_rows=16, grid_cols=16)
    model = L.alloc_RC_model(C.byref(cfg), flp2, None, None, 0, 0)
    _, msg = _expect_error(lib, lambda: L.populate_R_model(model, flp2))
    assert "inordinate floorplan size" in msg
    L.delete_RC_model(model)
    L.free_flp(flp, 0, 0)
    L.free_flp(flp2, 0, 0)
