"""CPU oracle for HotSpot's grid thermal model (numpy/scipy restatement).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import this module, and only as the
checker.  The product library never calls it.

The reference computes the grid model with per-cell loops (``find_res`` and
the NP/SP/.. macros, temperature_grid.c:26-136, :3096-3122).  This module
restates the same discrete model as an explicit sparse conductance matrix:

    C dT/dt = P + b - G T        (transient, temperature_grid.c:3130-3601)
    G T     = P + b              (steady,    temperature_grid.c:2426-2608)

where G holds every face conductance of the grid cells and the 12 (or 28)
package nodes, b holds the ambient/inlet sources and C the capacitances.
The steady solution is a direct sparse solve of that system (the fully
converged fixed point of the reference's Gauss-Seidel/multigrid iteration),
and the transient is the reference's adaptive RK4 (RCutil.c:440-640) applied
to ``f(T) = (P + b - G T) / C``.

How each part is pinned (see tests/test_oracle.py):
  * layer R/C, package R/C, block<->grid maps: golden vectors written by the
    reference library (tests/golden/), exact to 1e-12 relative;
  * the operator (G, b, C): the reference's own ``slope_fn_grid`` evaluated on
    seeded random temperatures, to 1e-10 relative;
  * the steady solve: the reference's own Gauss-Seidel sweep driven to a
    per-sweep change below 1e-13 K, and invariants (symmetry, energy balance,
    maximum principle, linearity);
  * the transient: the reference's own ``compute_temp`` trajectories;
  * backward Euler (``advance_be``): parity unpinned against the reference,
    whose SuperLU build is not available; pinned by its limits.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# --------------------------------------------------------------------------
# constants (temperature.h:29-119, util.h:18-19, temperature_grid.h:20-58,
# microchannel.h:8-14)
# --------------------------------------------------------------------------
DELTA = 1.0e-6
LARGENUM = 1.0e100
C_FACTOR = 0.333
MIN_STEP = 1e-7
EXTRA, EXTRA_SEC = 12, 16
(SP_W, SP_E, SP_N, SP_S, SINK_C_W, SINK_C_E, SINK_C_N, SINK_C_S,
 SINK_W, SINK_E, SINK_N, SINK_S) = range(12)
(SUB_W, SUB_E, SUB_N, SUB_S, SOLDER_W, SOLDER_E, SOLDER_N, SOLDER_S,
 PCB_C_W, PCB_C_E, PCB_C_N, PCB_C_S, PCB_W, PCB_E, PCB_N, PCB_S) = range(12, 28)
K_METAL, K_C4, K_SUB, K_SOLDER, K_PCB = 1 / 0.0025, 1 / 0.8, 1 / 0.5, 1 / 0.06, 1 / 0.333
SPEC_HEAT_METAL, SPEC_HEAT_C4, SPEC_HEAT_SUB = 3.55e6, 1.65e6, 1.6e6
SPEC_HEAT_SOLDER, SPEC_HEAT_PCB = 2.1e6, 1.32e6
OCCUPANCY_THRESHOLD = 0.95
TSV, SOLID, FLUID, INLET, OUTLET = -1, 0, 1, 2, 3
GRID_AVG, GRID_MIN, GRID_MAX, GRID_CENTER = 0, 1, 2, 3
MAP_MODES = {"avg": GRID_AVG, "min": GRID_MIN, "max": GRID_MAX, "center": GRID_CENTER}

# default_thermal_config (temperature.c:19-109)
DEFAULT_CONFIG = dict(
    t_chip=0.15e-3, k_chip=100.0, p_chip=1.75e6, thermal_threshold=81.8 + 273.15,
    c_convec=140.4, r_convec=0.1, s_sink=60e-3, t_sink=6.9e-3, k_sink=400.0,
    p_sink=3.55e6, s_spreader=30e-3, t_spreader=1e-3, k_spreader=400.0,
    p_spreader=3.55e6, t_interface=20e-6, k_interface=4.0, p_interface=4.0e6,
    model_secondary=0, r_convec_sec=1.0, c_convec_sec=140.4, n_metal=8,
    t_metal=10.0e-6, t_c4=0.0001, s_c4=20.0e-6, n_c4=400, s_sub=0.021,
    t_sub=0.001, s_solder=0.021, t_solder=0.00094, s_pcb=0.1, t_pcb=0.002,
    ambient=45 + 273.15, init_temp=60 + 273.15, sampling_intvl=3.333e-6,
    grid_rows=64, grid_cols=64, grid_map_mode="center", grid_layer_file=None,
    detailed_3D_used=0,
)

# default_microchannel_config (microchannel.c:17-43)
DEFAULT_MICROCHANNEL = dict(
    pumping_pressure=5000.0, pump_internal_res=0.0, inlet_temperature=300.0,
    coolant_capac=4172638.0, coolant_res=1.647717911, coolant_visc=0.000889,
    wall_capac=1635660.0, wall_res=0.0076923077, htc=27132.0,
)


def getr(k, t, area):
    """R = t / (k*A)  (RCutil.c:20-23)."""
    return t / (k * area)


def getcap(sp_heat, t, area):
    """C = C_FACTOR * c_v * t * A  (RCutil.c:26-31)."""
    return C_FACTOR * sp_heat * t * area


def eq(x, y):
    """util.c:17-20."""
    return abs(x - y) < DELTA


def tolerant_ceil(v):
    """util.c:57-66."""
    nearest = math.floor(v + 0.5)
    return int(nearest) if eq(v, nearest) else int(math.ceil(v))


def tolerant_floor(v):
    """util.c:68-77."""
    nearest = math.floor(v + 0.5)
    return int(nearest) if eq(v, nearest) else int(math.floor(v))


# --------------------------------------------------------------------------
# floorplans (flp.c:1004-1218, :334-350, :1501-1531)
# --------------------------------------------------------------------------
@dataclass
class Unit:
    name: str
    width: float
    height: float
    leftx: float
    bottomy: float
    specificheat: float = 0.0
    resistivity: float = 0.0
    hasRes: bool = False
    hasSh: bool = False


def parse_flp_text(text: str) -> list[Unit]:
    """Block lines ``name w h x y [c_v rho]``; connectivity lines are skipped
    (flp.c:1059-1118); the floorplan is moved to the origin (flp.c:1212)."""
    units = []
    for line in text.splitlines():
        tok = line.split()
        if not tok or tok[0].startswith("#"):
            continue
        nums = []
        for t in tok[1:]:
            try:
                nums.append(float(t))
            except ValueError:
                break
        if len(nums) >= 7 or len(nums) == 5:
            raise ValueError("invalid floorplan file format")
        if len(nums) == 6:
            units.append(Unit(tok[0], *nums[:4], specificheat=nums[4],
                              resistivity=nums[5], hasRes=True, hasSh=True))
        elif len(nums) == 4:
            units.append(Unit(tok[0], *nums[:4]))
        elif len(tok) == 3:
            continue  # connectivity information
        else:
            raise ValueError("invalid floorplan file format")
    if not units:
        raise ValueError("no units specified in the floorplan file")
    minx = min(u.leftx for u in units)
    miny = min(u.bottomy for u in units)
    for u in units:
        u.leftx += (0.0 - minx)
        u.bottomy += (0.0 - miny)
    return units


def read_flp(path: str) -> list[Unit]:
    with open(path) as f:
        return parse_flp_text(f.read())


def total_width(units):
    return max(u.leftx + u.width for u in units) - min(u.leftx for u in units)


def total_height(units):
    return max(u.bottomy + u.height for u in units) - min(u.bottomy for u in units)


# --------------------------------------------------------------------------
# layer stack (temperature_grid.c:459-893)
# --------------------------------------------------------------------------
@dataclass
class Layer:
    k: float
    thickness: float
    sp: float
    has_lateral: bool
    has_power: bool
    units: list
    map_id: int                     # layers sharing a block-grid map share map_id
    is_microchannel: bool = False
    mc: dict | None = None          # microchannel parameters for this layer
    rx: float = 0.0
    ry: float = 0.0
    rz: float = 0.0
    c: float = 0.0


def read_materials(path: str) -> dict:
    """materials.c:17-155: name, solid|fluid, k, c_v, [viscosity]."""
    toks = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if t and not t[0].startswith("#"):
                toks.append(t[0])
    out, i = {}, 0
    while i < len(toks):
        name, kind = toks[i], toks[i + 1]
        k, cv = float(toks[i + 2]), float(toks[i + 3])
        visc = 0.0
        i += 4
        if kind == "fluid":
            visc = float(toks[i])
            i += 1
        out[name] = dict(k=k, cv=cv, visc=visc)
    return out


def parse_lcf(path: str, materials: dict | None = None) -> list[dict]:
    """7-field layer records (temperature_grid.c:612-760): number, lateral
    Y/N, power Y/N, specific heat (or a material name, which replaces both
    specific heat and resistivity), resistivity, thickness, floorplan."""
    toks = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if t and not t[0].startswith("#"):
                toks.append(t[0])
    recs, i = [], 0
    base_dir = os.path.dirname(os.path.abspath(path))
    while i < len(toks):
        rec = dict(no=int(toks[i]), lateral=toks[i + 1][0] in "Yy",
                   power=toks[i + 2][0] in "Yy")
        try:
            rec["sp"] = float(toks[i + 3])
            rec["k"] = 1.0 / float(toks[i + 4])
            i += 5
        except ValueError:
            m = (materials or {})[toks[i + 3]]
            rec["sp"], rec["k"] = m["cv"], m["k"]
            i += 4
        rec["thickness"] = float(toks[i])
        rec["flp"] = toks[i + 1]
        i += 2
        recs.append(rec)
    return recs


def _microchannel_units(cell_types, cw, ch, mc):
    """Per-cell floorplan generated from the CSV network
    (microchannel.c:326-351), with every value rounded through '%e' exactly
    as the reference writes and re-reads it."""
    nr, nc = cell_types.shape
    r = lambda v: float("%e" % v)
    units = []
    for i in range(nr):
        for j in range(nc):
            fluid = cell_types[i, j] in (FLUID, INLET, OUTLET)
            cv = mc["coolant_capac"] if fluid else mc["wall_capac"]
            res = mc["coolant_res"] if fluid else mc["wall_res"]
            units.append(Unit(f"Cell_{i}_{j}", r(cw), r(ch), r(j * cw),
                              r((nr - i - 1) * ch), r(cv), r(res), True, True))
    minx = min(u.leftx for u in units)
    miny = min(u.bottomy for u in units)
    for u in units:
        u.leftx += -minx
        u.bottomy += -miny
    return units


def read_network_csv(path: str, nr: int, nc: int) -> np.ndarray:
    """microchannel.c:258-324: comma/space separated integer cell types."""
    types = np.zeros((nr, nc), dtype=np.int64)
    with open(path) as f:
        rows = [ln for ln in f.read().split("\n")]
    i = 0
    for ln in rows:
        cells = [c for c in re.split(r"[ ,]+", ln.strip()) if c != ""]
        if not cells:
            continue
        if len(cells) > nc or i >= nr:
            raise ValueError("microchannel network larger than the grid")
        types[i, :len(cells)] = [int(c) for c in cells]
        i += 1
    return types


def hydro_conductance(h, w, L, visc):
    """Rectangular-duct hydraulic conductance (microchannel.c:361-376)."""
    if h == w:
        return (0.42229 * h ** 4) / (12 * visc * L)
    if h > w:
        return ((1 - 0.63 * (w / h)) * w ** 3 * h) / (12 * visc * L)
    return ((1 - 0.63 * (h / w)) * h ** 3 * w) / (12 * visc * L)


def solve_pressure(cell_types: np.ndarray, pumping_pressure: float) -> np.ndarray:
    """Pressure of every fluid cell for an ideal pump (microchannel.c:378-565,
    solved there by dense Gauss-Jordan, util.c:621-699): Laplace equations on
    FLUID cells over their fluid neighbours, p = pumping_pressure at inlets,
    p = 0 at outlets.  Returns an nr x nc array (0 for solid cells)."""
    nr, nc = cell_types.shape
    fluid = np.isin(cell_types, (FLUID, INLET, OUTLET))
    idx = -np.ones((nr, nc), dtype=np.int64)
    idx[fluid] = np.arange(fluid.sum())
    n = int(fluid.sum())
    rows, cols, vals = [], [], []
    b = np.zeros(n)
    for i in range(nr):
        for j in range(nc):
            if not fluid[i, j]:
                continue
            r = idx[i, j]
            t = cell_types[i, j]
            if t == FLUID:
                diag = 0.0
                for (a, c) in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                    if 0 <= a < nr and 0 <= c < nc and fluid[a, c]:
                        rows.append(r); cols.append(idx[a, c]); vals.append(-1.0)
                        diag += 1.0
                rows.append(r); cols.append(r); vals.append(diag)
            else:
                rows.append(r); cols.append(r); vals.append(1.0)
                if t == INLET:
                    b[r] = pumping_pressure
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    p = spla.spsolve(A.tocsc(), b)
    out = np.zeros((nr, nc))
    out[fluid] = p
    return out


# --------------------------------------------------------------------------
# the model
# --------------------------------------------------------------------------
class OracleModel:
    """Grid thermal model built from the same inputs as the reference:
    thermal config values, a default floorplan file or a layer configuration
    file, and the detailed-3D / microchannel switches."""

    def __init__(self, config: dict | None = None, flp: str | None = None,
                 lcf: str | None = None, detailed_3D=False, use_microchannels=False,
                 microchannel: dict | None = None, materials: dict | None = None,
                 superlu_outlets=False):
        # superlu_outlets: outlet cells' advection as the SuperLU build's
        # matrix has it (build_transient_grid_matrix, temperature_grid.c:
        # 5313-5345) instead of the slope's (:3256-3590)
        self.superlu_outlets = bool(superlu_outlets)
        cfg = dict(DEFAULT_CONFIG)
        cfg.update(config or {})
        if lcf is not None:
            cfg["grid_layer_file"] = lcf
        self.cfg = cfg
        # any grid size, as the reference's SuperLU build takes (its
        # iterative build asks for powers of two, temperature_grid.c:901-905)
        self.nr, self.nc = int(cfg["grid_rows"]), int(cfg["grid_cols"])
        if self.nr <= 0 or self.nc <= 0:
            raise ValueError("grid rows and columns must be positive")
        self.secondary = bool(cfg["model_secondary"])
        self.detailed = bool(detailed_3D or cfg.get("detailed_3D_used", 0))
        self.use_mc = bool(use_microchannels)
        self.mc_default = dict(DEFAULT_MICROCHANNEL)
        self.mc_default.update(microchannel or {})
        self.map_mode = MAP_MODES[cfg["grid_map_mode"]]
        self.nx = EXTRA + (EXTRA_SEC if self.secondary else 0)
        self.has_lcf = cfg.get("grid_layer_file") is not None
        if self.detailed and not self.has_lcf:
            raise ValueError("detailed 3D modeling needs a layer configuration file")
        if self.use_mc and not self.detailed:
            raise ValueError("microchannels need detailed 3D modeling")
        self._build_layers(flp, materials)
        self._populate_R()
        self._populate_C()
        self._assemble()

    # ---- stack ---------------------------------------------------------
    def _build_layers(self, flp_path, materials):
        """populate_layers_grid (:835-893) and friends."""
        cfg = self.cfg
        L = []
        if self.has_lcf:
            recs = parse_lcf(cfg["grid_layer_file"], materials)
            base_dir = os.path.dirname(os.path.abspath(cfg["grid_layer_file"]))
            width = height = None
            chip = [None] * len(recs)
            for mid, r in enumerate(recs):
                fpath = r["flp"] if os.path.isabs(r["flp"]) else os.path.join(base_dir, r["flp"])
                if not os.path.exists(fpath) and os.path.exists(r["flp"]):
                    fpath = r["flp"]
                lay = Layer(r["k"], r["thickness"], r["sp"], r["lateral"], r["power"],
                            None, mid)
                if fpath.endswith(".csv") and self.use_mc and width is None:
                    raise ValueError("a microchannel layer cannot be the first layer")
                if fpath.endswith(".csv") and self.use_mc:
                    # cell size uses the width fixed by the first layer (:708-716)
                    cw = (width or 0.0) / self.nc
                    ch = (height or 0.0) / self.nr
                    types = read_network_csv(fpath, self.nr, self.nc)
                    mc = dict(self.mc_default)
                    mc.update(cell_width=cw, cell_height=ch, cell_thickness=r["thickness"],
                              cell_types=types)
                    mc["pressure"] = solve_pressure(types, mc["pumping_pressure"])
                    mc["hydroC"] = hydro_conductance(r["thickness"], cw, ch, mc["coolant_visc"])
                    lay.units = _microchannel_units(types, cw, ch, mc)
                    lay.is_microchannel = True
                    lay.mc = mc
                elif fpath.endswith(".csv"):
                    raise ValueError("microchannel floorplan but use_microchannels = 0")
                else:
                    lay.units = read_flp(fpath)
                if width is None:
                    width, height = total_width(lay.units), total_height(lay.units)
                elif not (eq(width, total_width(lay.units)) and eq(height, total_height(lay.units))):
                    raise ValueError("width and height differ across layers")
                chip[r["no"]] = lay
            self.width, self.height = width, height
        else:
            units = read_flp(flp_path)
            si = Layer(cfg["k_chip"], cfg["t_chip"], cfg["p_chip"], True, True, units, 0)
            tim = Layer(cfg["k_interface"], cfg["t_interface"], cfg["p_interface"],
                        True, False, units, 0)
            chip = [si, tim]
            if self.secondary:
                c4 = Layer(K_C4, cfg["t_c4"], SPEC_HEAT_C4, True, False, units, 0)
                metal = Layer(K_METAL, cfg["t_metal"], SPEC_HEAT_METAL, True, False, units, 0)
                chip = [c4, metal, si, tim]
            self.width, self.height = total_width(units), total_height(units)
        last = chip[-1]
        pkg_map = last.map_id if not self.detailed else 1000
        spreader = Layer(cfg["k_spreader"], cfg["t_spreader"], cfg["p_spreader"], True,
                         False, last.units, pkg_map)
        sink = Layer(cfg["k_sink"], cfg["t_sink"], cfg["p_sink"], True, False,
                     last.units, pkg_map)
        L = chip + [spreader, sink]
        if self.secondary:
            # secondary package layers 0-2 take the floorplan of the topmost
            # chip layer (append_package_layers, :566-606)
            top = chip[0] if self.has_lcf else chip[2]
            sec_map = top.map_id if not self.detailed else 1001
            pcb = Layer(K_PCB, cfg["t_pcb"], SPEC_HEAT_PCB, True, False, top.units, sec_map)
            solder = Layer(K_SOLDER, cfg["t_solder"], SPEC_HEAT_SOLDER, True, False, top.units, sec_map)
            sub = Layer(K_SUB, cfg["t_sub"], SPEC_HEAT_SUB, True, False, top.units, sec_map)
            L = [pcb, solder, sub] + L
        self.layers = L
        self.nl = len(L)
        self.spidx, self.hsidx = self.nl - 2, self.nl - 1
        if self.secondary:
            self.subidx, self.solderidx, self.pcbidx = 2, 1, 0
        if (self.width > cfg["s_sink"] or self.height > cfg["s_sink"]
                or self.width > cfg["s_spreader"] or self.height > cfg["s_spreader"]):
            raise ValueError("inordinate floorplan size!")
        self.cw = self.width / self.nc
        self.ch = self.height / self.nr
        self.total_n_blocks = sum(len(l.units) for l in L)
        self._build_maps()

    # ---- block-grid maps (set_bgmap :341-456, new_blist :193-224,
    #      blist_append :286-320) ----------------------------------------
    def _build_maps(self):
        nr, nc, cw, ch, H = self.nr, self.nc, self.cw, self.ch, self.height
        self.maps = {}
        for n, lay in enumerate(self.layers):
            if lay.map_id in self.maps:
                continue
            entries = [[[] for _ in range(nc)] for _ in range(nr)]
            g2b = []
            # detailed per-cell (rx, ry, rz, cap, lock, hasRes) of the list head
            det = {}
            for u, un in enumerate(lay.units):
                lu, ru = un.leftx, un.leftx + un.width
                bu, tu = un.bottomy, un.bottomy + un.height
                i1 = nr - tolerant_ceil(tu / ch)
                i2 = nr - tolerant_floor(bu / ch)
                j1 = tolerant_floor(lu / cw)
                j2 = tolerant_ceil(ru / cw)
                if i1 < 0 or j1 < 0:
                    raise ValueError("negative grid cell start index!")
                if i2 > nr or j2 > nc:
                    raise ValueError("grid cell end index out of bounds!")
                if i1 >= i2 or j1 >= j2:
                    raise ValueError("invalid floorplan spec or grid resolution")
                g2b.append((i1, i2, j1, j2))
                res = un.resistivity if (un.hasRes and self.detailed) else 1.0 / lay.k
                sh = un.specificheat if (un.hasSh and self.detailed) else lay.sp
                t = lay.thickness
                for i in range(i1, i2):
                    for j in range(j1, j2):
                        if i > i1 and i < i2 - 1 and j > j1 and j < j2 - 1:
                            occ = 1.0
                            first_flag = 1
                        else:
                            lc, rc = j * cw, (j + 1) * cw
                            tc, bc = H - i * ch, H - (i + 1) * ch
                            oh = min(tu, tc) - max(bu, bc)
                            ow = min(ru, rc) - max(lu, lc)
                            if eq(oh / ch, 0):
                                oh = 0.0
                            elif eq(oh / ch, 1):
                                oh = ch
                            if eq(ow / cw, 0):
                                ow = 0.0
                            elif eq(ow / cw, 1):
                                ow = cw
                            occ = (oh * ow) / (ch * cw)
                            if oh < 0 or ow < 0:
                                raise ValueError("negative overlap!")
                            first_flag = 0
                        lst = entries[i][j]
                        if self.detailed:
                            gx = 1.0 / getr(1.0 / res, cw, ch * t)
                            gy = 1.0 / getr(1.0 / res, ch, cw * t)
                            gz = 1.0 / getr(1.0 / res, t, cw * ch)
                            if not lst:
                                if occ >= OCCUPANCY_THRESHOLD:
                                    det[(i, j)] = [getr(1 / res, cw, ch * t), getr(1 / res, ch, cw * t),
                                                   getr(1 / res, t, cw * ch), getcap(sh, t, cw * ch),
                                                   True, un.hasSh]
                                else:
                                    det[(i, j)] = [1 / (gx * occ), 1 / (gy * occ), 1 / (gz * occ),
                                                   getcap(sh, t, cw * ch), False, un.hasSh]
                            else:
                                d = det[(i, j)]
                                if not d[4]:
                                    if occ >= OCCUPANCY_THRESHOLD:
                                        d[0] = getr(1 / res, cw, ch * t)
                                        d[1] = getr(1 / res, ch, cw * t)
                                        d[2] = getr(1 / res, t, cw * ch)
                                        d[3] = getcap(sh, t, cw * ch)
                                        d[4] = True
                                    else:
                                        d[0] = 1 / ((1 / d[0]) + gx * occ)
                                        d[1] = 1 / ((1 / d[1]) + gy * occ)
                                        d[2] = 1 / ((1 / d[2]) + gz * occ)
                                        d[4] = False
                        lst.append((u, occ))
            ec, eu, eo = [], [], []
            for i in range(nr):
                for j in range(nc):
                    if not entries[i][j]:
                        raise ValueError(f"grid cell ({i},{j}) not covered by the floorplan")
                    for (u, occ) in entries[i][j]:
                        ec.append(i * nc + j); eu.append(u); eo.append(occ)
            self.maps[lay.map_id] = dict(
                entries=entries, g2b=np.array(g2b, dtype=np.int64), det=det,
                cell=np.array(ec, dtype=np.int64), unit=np.array(eu, dtype=np.int64),
                occ=np.array(eo))

    # ---- R and C (populate_R_model_grid :950-1063, populate_C :1065-1114,
    #      populate_package_R/C temperature.c:458-569) -------------------
    def _populate_R(self):
        cfg, cw, ch = self.cfg, self.cw, self.ch
        for n, lay in enumerate(self.layers):
            if lay.has_lateral:
                lay.rx = getr(lay.k, cw, ch * lay.thickness)
                lay.ry = getr(lay.k, ch, cw * lay.thickness)
            else:
                lay.rx = lay.ry = LARGENUM
            lay.rz = getr(lay.k, lay.thickness, cw * ch)
            if n == self.hsidx:
                lay.rz += cfg["r_convec"] * (cfg["s_sink"] * cfg["s_sink"]) / (cw * ch)
        self.pack = package_R(cfg, self.width, self.height)

    def _populate_C(self):
        cfg, cw, ch = self.cfg, self.cw, self.ch
        for lay in self.layers:
            lay.c = getcap(lay.sp, lay.thickness, cw * ch)
        self.layers[self.hsidx].c += C_FACTOR * cfg["c_convec"] * (cw * ch) / (cfg["s_sink"] * cfg["s_sink"])
        if self.secondary:
            self.layers[self.pcbidx].c += C_FACTOR * cfg["c_convec_sec"] * (cw * ch) / (cfg["s_pcb"] * cfg["s_pcb"])
        self.pack.update(package_C(cfg, self.width, self.height))

    # ---- per-cell detailed values (find_res_3D :144-173, find_cap_3D :176-186)
    def _cell_arrays(self, n):
        lay = self.layers[n]
        nr, nc = self.nr, self.nc
        rx = np.full((nr, nc), lay.rx)
        ry = np.full((nr, nc), lay.ry)
        rz = np.full((nr, nc), lay.rz)
        cap = np.full((nr, nc), lay.c)
        if self.detailed:
            pkg = n in (self.spidx, self.hsidx) or (self.secondary and n <= 2)
            if not pkg:
                det = self.maps[lay.map_id]["det"]
                for (i, j), d in det.items():
                    rx[i, j], ry[i, j], rz[i, j] = d[0], d[1], d[2]
                    if d[4]:
                        cap[i, j] = d[3]
        return rx, ry, rz, cap

    def _fluid(self, n):
        lay = self.layers[n]
        if not lay.is_microchannel:
            return np.zeros((self.nr, self.nc), dtype=bool)
        return np.isin(lay.mc["cell_types"], (FLUID, INLET, OUTLET))

    # ---- operator assembly ---------------------------------------------
    def _assemble(self):
        nl, nr, nc, nx = self.nl, self.nr, self.nc, self.nx
        ncell = nl * nr * nc
        N = ncell + nx
        self.N, self.ncell = N, ncell
        cw, ch, cfg, pk = self.cw, self.ch, self.cfg, self.pack
        L = self.layers
        idx = np.arange(ncell).reshape(nl, nr, nc)
        rows, cols, vals = [], [], []      # symmetric couplings g between a,b
        diag = np.zeros(N)                 # extra diagonal (ambient) terms
        b = np.zeros(N)
        amb = cfg["ambient"]

        def couple(a, c, g):
            a, c, g = np.broadcast_arrays(np.asarray(a), np.asarray(c),
                                          np.asarray(g, dtype=float))
            rows.append(a.ravel()); cols.append(c.ravel()); vals.append(g.ravel())

        def ground(a, g):
            a = np.asarray(a).ravel()
            g = np.broadcast_to(np.asarray(g, dtype=float), a.shape).ravel()
            np.add.at(diag, a, g)
            np.add.at(b, a, g * amb)

        cells = [self._cell_arrays(n) for n in range(nl)]
        fluid = [self._fluid(n) for n in range(nl)]
        self.cap_cells = np.stack([c[3] for c in cells])
        # lateral faces (find_res :26-136)
        for n in range(nl):
            rx, ry, rz, _ = cells[n]
            fl = fluid[n]
            # row neighbours (i, i+1): rx/2 + rx/2
            rr = rx[:-1, :] / 2.0 + rx[1:, :] / 2.0
            cc = ry[:, :-1] / 2.0 + ry[:, 1:] / 2.0
            if L[n].is_microchannel:
                htc, t = L[n].mc["htc"], L[n].thickness
                mixed_r = fl[:-1, :] != fl[1:, :]
                solid_r = np.where(fl[:-1, :], rx[1:, :], rx[:-1, :])
                hr = solid_r / 2.0 + 1.0 / (htc * cw * t)
                mixed_c = fl[:, :-1] != fl[:, 1:]
                solid_c = np.where(fl[:, :-1], ry[:, 1:], ry[:, :-1])
                hc = solid_c / 2.0 + 1.0 / (htc * ch * t)
                rr = np.where(mixed_r, hr, rr)
                cc = np.where(mixed_c, hc, cc)
            couple(idx[n, :-1, :], idx[n, 1:, :], 1.0 / rr)
            couple(idx[n, :, :-1], idx[n, :, 1:], 1.0 / cc)
        # vertical faces n <-> n+1
        for n in range(nl - 1):
            a_rz, b_rz = cells[n][2], cells[n + 1][2]
            fa, fb = fluid[n], fluid[n + 1]
            if self.detailed or self.use_mc:
                if n == 0:
                    r = a_rz + b_rz / 2.0
                else:
                    r = a_rz / 2.0 + b_rz / 2.0
                mixed = fa != fb
                if mixed.any():
                    htc = (L[n + 1].mc if L[n + 1].is_microchannel else L[n].mc)["htc"]
                    if n == 0:
                        hr = a_rz + 1.0 / (htc * ch * cw)
                    else:
                        hr = np.where(fa, b_rz, a_rz) / 2.0 + 1.0 / (htc * ch * cw)
                    r = np.where(mixed, hr, r)
                if not self.detailed:
                    r = np.where(mixed, r, L[n].rz)
            else:
                r = np.full((nr, nc), L[n].rz)
            couple(idx[n], idx[n + 1], 1.0 / r)
        # ambient: sink cells (:2501-2503) and PCB cells (:2556-2559)
        hs = self.hsidx
        ground(idx[hs], 1.0 / L[hs].rz)
        if self.secondary:
            ground(idx[self.pcbidx], 1.0 / (cfg["r_convec_sec"] * (cfg["s_pcb"] * cfg["s_pcb"]) / (cw * ch)))
        X = ncell  # first package node

        def ring(n, node_n, node_s, node_e, node_w, r1_y, r1_x):
            lay = L[n]
            gy = 1.0 / (lay.ry / 2.0 + nc * r1_y)
            gx = 1.0 / (lay.rx / 2.0 + nr * r1_x)
            couple(idx[n, 0, :], X + node_n, gy)
            couple(idx[n, nr - 1, :], X + node_s, gy)
            couple(idx[n, :, nc - 1], X + node_e, gx)
            couple(idx[n, :, 0], X + node_w, gx)

        ring(self.spidx, SP_N, SP_S, SP_E, SP_W, pk["r_sp1_y"], pk["r_sp1_x"])
        ring(hs, SINK_C_N, SINK_C_S, SINK_C_E, SINK_C_W, pk["r_hs1_y"], pk["r_hs1_x"])
        # package-node couplings (single_iteration_steady_pack :2090-2393)
        for o, c_ in ((SINK_N, SINK_C_N), (SINK_S, SINK_C_S)):
            ground(X + o, 1.0 / (pk["r_hs_per"] + pk["r_amb_per"]))
            couple(X + o, X + c_, 1.0 / (pk["r_hs2_y"] + pk["r_hs"]))
        for o, c_ in ((SINK_W, SINK_C_W), (SINK_E, SINK_C_E)):
            ground(X + o, 1.0 / (pk["r_hs_per"] + pk["r_amb_per"]))
            couple(X + o, X + c_, 1.0 / (pk["r_hs2_x"] + pk["r_hs"]))
        for c_, s_ in ((SINK_C_N, SP_N), (SINK_C_S, SP_S)):
            ground(X + c_, 1.0 / (pk["r_hs_c_per_y"] + pk["r_amb_c_per_y"]))
            couple(X + c_, X + s_, 1.0 / pk["r_sp_per_y"])
        for c_, s_ in ((SINK_C_W, SP_W), (SINK_C_E, SP_E)):
            ground(X + c_, 1.0 / (pk["r_hs_c_per_x"] + pk["r_amb_c_per_x"]))
            couple(X + c_, X + s_, 1.0 / pk["r_sp_per_x"])
        if self.secondary:
            ring(self.subidx, SUB_N, SUB_S, SUB_E, SUB_W, pk["r_sub1_y"], pk["r_sub1_x"])
            ring(self.solderidx, SOLDER_N, SOLDER_S, SOLDER_E, SOLDER_W,
                 pk["r_solder1_y"], pk["r_solder1_x"])
            ring(self.pcbidx, PCB_C_N, PCB_C_S, PCB_C_E, PCB_C_W, pk["r_pcb1_y"], pk["r_pcb1_x"])
            for o, c_ in ((PCB_N, PCB_C_N), (PCB_S, PCB_C_S)):
                ground(X + o, 1.0 / pk["r_amb_sec_per"])
                couple(X + o, X + c_, 1.0 / (pk["r_pcb2_y"] + pk["r_pcb"]))
            for o, c_ in ((PCB_W, PCB_C_W), (PCB_E, PCB_C_E)):
                ground(X + o, 1.0 / pk["r_amb_sec_per"])
                couple(X + o, X + c_, 1.0 / (pk["r_pcb2_x"] + pk["r_pcb"]))
            for c_, s_, sb in ((PCB_C_N, SOLDER_N, SUB_N), (PCB_C_S, SOLDER_S, SUB_S)):
                ground(X + c_, 1.0 / pk["r_amb_sec_c_per_y"])
                couple(X + c_, X + s_, 1.0 / pk["r_pcb_c_per_y"])
                couple(X + s_, X + sb, 1.0 / pk["r_solder_per_y"])
            for c_, s_, sb in ((PCB_C_W, SOLDER_W, SUB_W), (PCB_C_E, SOLDER_E, SUB_E)):
                ground(X + c_, 1.0 / pk["r_amb_sec_c_per_x"])
                couple(X + c_, X + s_, 1.0 / pk["r_pcb_c_per_x"])
                couple(X + s_, X + sb, 1.0 / pk["r_solder_per_x"])
        r = np.concatenate(rows); c = np.concatenate(cols); v = np.concatenate(vals)
        off = sp.coo_matrix((v, (r, c)), shape=(N, N))
        off = off + off.T
        deg = np.asarray(off.sum(axis=1)).ravel()
        G = sp.diags(deg + diag) - off
        # microchannel advection (slope_fn_grid :3256-3590): non-symmetric
        adv_r, adv_c, adv_v = [], [], []
        for n in range(nl):
            if not L[n].is_microchannel:
                continue
            mc = L[n].mc
            p, hC, Cv = mc["pressure"], mc["hydroC"], mc["coolant_capac"]
            fl = fluid[n]
            types = mc["cell_types"]
            for i in range(nr):
                for j in range(nc):
                    if not fl[i, j]:
                        continue
                    me = idx[n, i, j]
                    if self.superlu_outlets and types[i, j] == OUTLET:
                        # the first fluid neighbour alone (N, S, W, E): -Cv Q/2
                        # to it, +Cv Q/2 on the diagonal (Jn/Js/Jw/Je, Jc)
                        for (a, cc_) in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                            if 0 <= a < nr and 0 <= cc_ < nc and fl[a, cc_]:
                                coef = Cv * (p[a, cc_] - p[i, j]) * hC
                                adv_r += [me, me]; adv_c += [idx[n, a, cc_], me]
                                adv_v += [-coef / 2.0, coef / 2.0]
                                break
                        continue
                    inflow = 0.0
                    for (a, cc_) in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                        if 0 <= a < nr and 0 <= cc_ < nc and fl[a, cc_]:
                            q = (p[a, cc_] - p[i, j]) * hC       # flow nbr -> me
                            coef = Cv * q
                            adv_r += [me, me]; adv_c += [idx[n, a, cc_], me]
                            adv_v += [-coef / 2.0, -coef / 2.0]
                            inflow += q
                    on_edge = i == 0 or i == nr - 1 or j == 0 or j == nc - 1
                    if types[i, j] == INLET and on_edge:
                        b[me] += Cv * (-inflow) * mc["inlet_temperature"]
                    elif types[i, j] == OUTLET and on_edge:
                        adv_r.append(me); adv_c.append(me); adv_v.append(Cv * inflow)
        if adv_r:
            G = G + sp.coo_matrix((adv_v, (adv_r, adv_c)), shape=(N, N))
        self.G = G.tocsr()
        self.b = b
        self.symmetric = not adv_r
        # The transient package equations of the solder nodes leave out their
        # coupling to the substrate nodes, which the steady equations include
        # (slope_fn_pack :3034-3070 vs single_iteration_steady_pack
        # :2314-2352), so the transient operator differs from the steady one
        # in those four rows.
        Gt = self.G.tolil(copy=True)
        if self.secondary:
            for s_, sb in ((SOLDER_N, SUB_N), (SOLDER_S, SUB_S)):
                g = 1.0 / pk["r_solder_per_y"]
                Gt[X + s_, X + s_] -= g
                Gt[X + s_, X + sb] += g
            for s_, sb in ((SOLDER_W, SUB_W), (SOLDER_E, SUB_E)):
                g = 1.0 / pk["r_solder_per_x"]
                Gt[X + s_, X + s_] -= g
                Gt[X + s_, X + sb] += g
        self.Gt = Gt.tocsr()
        Cvec = np.zeros(N)
        Cvec[:ncell] = self.cap_cells.ravel()
        Cvec[ncell:] = package_node_caps(pk, self.secondary)
        self.Cvec = Cvec

    # ---- vectors -------------------------------------------------------
    def n_vector(self):
        return self.total_n_blocks + self.nx

    def layer_bases(self):
        out, base = [], 0
        for lay in self.layers:
            out.append(base)
            base += len(lay.units)
        return out

    def b2g(self, vec, power=True):
        """xlate_vector_b2g (:1888-1929) with blist_avg (:323-338)."""
        nl, nr, nc = self.nl, self.nr, self.nc
        area = (self.width * self.height) / (nc * nr)
        out = np.zeros(self.N)
        g = out[:self.ncell].reshape(nl, nr, nc)
        vec = np.asarray(vec, dtype=float)
        for n, base in enumerate(self.layer_bases()):
            lay = self.layers[n]
            m = self.maps[lay.map_id]
            v = vec[base + m["unit"]]
            term = m["occ"] * v
            if power:
                a = np.array([un.width * un.height for un in lay.units])
                term = term / a[m["unit"]]
            acc = np.zeros(nr * nc)
            np.add.at(acc, m["cell"], term)   # sequential: list order per cell
            g[n] = (acc * area if power else acc).reshape(nr, nc)
        out[self.ncell:] = vec[self.total_n_blocks:self.total_n_blocks + self.nx]
        return out

    def power_grid(self, block_power):
        """Grid power with the package-node power zeroed
        (set_internal_power_grid :2018-2024)."""
        p = np.array(block_power, dtype=float, copy=True)
        p[self.total_n_blocks:] = 0.0
        return self.b2g(p, power=True)

    def g2b(self, gvec):
        """xlate_temp_g2b (:1932-2015)."""
        nl, nr, nc = self.nl, self.nr, self.nc
        g = gvec[:self.ncell].reshape(nl, nr, nc)
        out = np.zeros(self.n_vector())
        for n, base in enumerate(self.layer_bases()):
            lay = self.layers[n]
            g2b = self.maps[lay.map_id]["g2b"]
            for u in range(len(lay.units)):
                i1, i2, j1, j2 = (int(x) for x in g2b[u])
                if self.map_mode == GRID_CENTER:
                    ci1, cj1 = (i1 + i2) // 2, (j1 + j2) // 2
                    ci2 = ci1 - (1 if (i2 - i1) % 2 == 0 else 0)
                    cj2 = cj1 - (1 if (j2 - j1) % 2 == 0 else 0)
                    out[base + u] = (g[n, ci1, cj1] + g[n, ci2, cj1] + g[n, ci1, cj2] + g[n, ci2, cj2]) / 4
                    continue
                blk = g[n, i1:i2, j1:j2]
                if self.map_mode == GRID_AVG:
                    s = 0.0
                    for v in blk.ravel():
                        s += v
                    out[base + u] = s / blk.size
                elif self.map_mode == GRID_MIN:
                    out[base + u] = blk.min()
                else:
                    out[base + u] = blk.max()
        out[self.total_n_blocks:] = gvec[self.ncell:self.ncell + self.nx]
        return out

    # ---- operator-level queries ----------------------------------------
    def slope(self, T, Pgrid):
        """dT/dt = (P + b - Gt T) / C  (slope_fn_grid + slope_fn_pack)."""
        return (Pgrid + self.b - self.Gt @ T) / self.Cvec

    def residual(self, T, Pgrid):
        return Pgrid + self.b - self.G @ T

    def steady_grid(self, Pgrid):
        """Converged steady state of the reference operator: a sparse direct
        solve, or above 150 000 nodes of a symmetric operator (400 000 of a
        non-symmetric one) Jacobi-preconditioned CG (BiCGStab) to a 1e-14
        relative residual, where the direct solve's fill-in takes minutes."""
        if self.N <= (150_000 if self.symmetric else 400_000):
            return spla.spsolve(self.G.tocsc(), Pgrid + self.b)
        return self._cg(Pgrid + self.b)

    def _cg(self, rhs, tol=1e-14, maxiter=200000):
        d = self.G.diagonal()
        M = sp.diags(1.0 / d)
        x0 = np.full(self.N, self.cfg["ambient"])
        if self.symmetric:
            x, info = spla.cg(self.G, rhs, x0=x0, rtol=tol, maxiter=maxiter, M=M)
        else:
            x, info = spla.bicgstab(self.G, rhs, x0=x0, rtol=tol, maxiter=maxiter, M=M)
        if info != 0:
            raise RuntimeError(f"oracle iterative solve did not converge ({info})")
        return x

    def steady(self, block_power):
        """steady_state_temp: block power -> block temperatures, plus the grid."""
        Pg = self.power_grid(block_power)
        Tg = self.steady_grid(Pg)
        return self.g2b(Tg), Tg

    # ---- transient: adaptive RK4 (RCutil.c:440-640) ---------------------
    def rk4_core(self, y, k1, Pg, h):
        f = lambda v: self.slope(v, Pg)
        t = y + h / 2.0 * k1
        k2 = f(t)
        t = y + h / 2.0 * k2
        k3 = f(t)
        t = y + h * k3
        k4 = f(t)
        return y + h * (k1 + 2 * k2 + 2 * k3 + k4) / 6.0

    def rk4(self, y, Pg, h):
        """One adaptive step with step doubling.  Returns (y_new, h_used,
        new_h).  Keeps the reference's controller exactly: k1 is replaced by
        the slope at the half step and reused by a retry, and the accepted
        value is the single full step (RCutil.c:568, :624-630)."""
        k1 = self.slope(y, Pg)
        new_h = h
        while True:
            h = new_h
            ytemp = self.rk4_core(y, k1, Pg, h)
            t1 = self.rk4_core(y, k1, Pg, h / 2.0)
            k1 = self.slope(t1, Pg)
            t2 = self.rk4_core(t1, k1, Pg, h / 2.0)
            mx = float(np.max(np.abs(ytemp - t2)))
            if mx <= 0.01:
                new_h = 0.95 * h * (abs(0.01 / mx) ** 0.2 if mx > 0 else math.inf)
                if new_h > 5.0 * h:
                    new_h = 5.0 * h
            else:
                new_h = 0.95 * h * abs(0.01 / mx) ** 0.25
                if new_h < h / 10.0:
                    new_h = h / 10.0
            if not new_h < h:
                break
        return ytemp, h, new_h

    def advance_exact(self, Tg, Pg, dt):
        """Exact solution of the interval's linear ODE C dT/dt = Pg + b - Gt T:
        T(dt) = T_inf + exp(-C^-1 Gt dt) (T - T_inf), T_inf = Gt^-1 (Pg + b)
        (scipy's expm_multiply).  The reference's adaptive RK4 accepts up to
        0.01 K of step-doubling error per step, so this is the yardstick for
        integrators that are more accurate than it."""
        rhs = Pg + self.b
        tinf = spla.spsolve(self.Gt.tocsc(), rhs)
        A = sp.diags(1.0 / self.Cvec) @ self.Gt
        return tinf + spla.expm_multiply(-dt * A, Tg - tinf)

    def advance_be(self, Tg, Pg, h):
        """One backward-Euler step of the whole interval, as compute_temp_grid
        does in the reference's SuperLU build (:3632-3662 calling
        backward_euler, RCutil.c:268-428): (C/h + G) T' = C T / h + Pg + b.
        G is the matrix that build_transient_grid_matrix assembles, whose
        secondary package links are the steady ones (the solder-substrate
        link in both directions, :6354 and :6440), and the secondary package
        nodes' capacitance is 0 (build_diagonal_matrix :6750-6768).  Parity
        unpinned against the reference (SuperLU absent); pinned by its limits
        in tests/test_oracle.py: h -> infinity gives the steady state, and the
        error against the exact solution falls linearly with h."""
        cbe = self.Cvec.copy()
        cbe[self.ncell + EXTRA:] = 0.0
        A = (sp.diags(cbe / h) + self.G).tocsc()
        return spla.spsolve(A, cbe / h * Tg + Pg + self.b)

    def advance(self, Tg, Pg, dt):
        """compute_temp_grid's integration loop (:3676-3692).  Returns the new
        grid state and the number of RK4 calls."""
        t, new_h, calls = 0.0, MIN_STEP, 0
        y = Tg
        while t < dt and new_h >= MIN_STEP * DELTA:
            h = new_h
            y, h, new_h = self.rk4(y, Pg, h)
            new_h = min(new_h, dt - t - h)
            calls += 1
            t += h
        return y, calls


# --------------------------------------------------------------------------
# the block model (temperature_block.c)
# --------------------------------------------------------------------------
def _touch(a1, a2, b1, b2):
    """The closed intervals [a1, a2] and [b1, b2] meet (flp.c:1431-1433)."""
    return (a1 <= b1 <= a2) or (a1 <= b2 <= a2) or (b1 <= a1 <= b2) or (b1 <= a2 <= b2)


def unit_neighbours(u, v):
    """How two floorplan units border each other (flp.c:1404-1499): returns
    ("x", shared length) for units side by side (a common vertical edge),
    ("y", shared length) for one above the other, or (None, 0).  Units that
    touch at a corner only are not neighbours.  All comparisons of edges use
    the floorplan tolerance eq() (1e-6 m)."""
    x1, x2, y1, y2 = u.leftx, u.leftx + u.width, u.bottomy, u.bottomy + u.height
    x3, x4, y3, y4 = v.leftx, v.leftx + v.width, v.bottomy, v.bottomy + v.height
    corner = ((eq(x2, x3) and eq(y2, y3)) or (eq(x1, x4) and eq(y1, y4)) or
              (eq(x2, x3) and eq(y1, y4)) or (eq(x1, x4) and eq(y2, y3)))
    if corner:
        return None, 0.0
    side = (eq(x1, x4) or eq(x2, x3)) and _touch(y1, y2, y3, y4)
    above = (eq(y1, y4) or eq(y2, y3)) and _touch(x1, x2, x3, x4)
    if above:       # get_shared_len takes the vertical case last
        return ("x" if side else "y"), min(x2, x4) - max(x1, x3)
    if side:
        return "x", min(y2, y4) - max(y1, y3)
    return None, 0.0


class OracleBlockModel:
    """HotSpot's block model (temperature_block.c) as a dense linear system.

    Nodes: four per floorplan unit, silicon/interface/spreader/sink (node
    layer*n + unit), then the 12 package nodes at 4n + k (SP_*, SINK_C_*,
    SINK_*, temperature.h:32-77).  Every link is a conductance between two
    nodes; the reference stores each link as two half-link conductances and
    combines them in series (populate_R_model_block, temperature_block.c:
    64-296).  Here each link's conductance is written directly:

      lateral, layer l, units i|j sharing an edge of length L:
          g_i = L * k_l t_l h_i / (w_i / 2) / h_i   (side by side, per unit
          width and height), series of g_i and g_j
      vertical, unit i: 1/r with r = t / (k A) of the upper layer
      spreader (sink) unit on the chip's north edge to SP_N (SINK_C_N):
          1 / (1/g_y + r_sp1_y g_N / g_y), g_N = sum of the edge units' g_y
          (likewise S with g_y, E and W with g_x)
      SP_d - SINK_C_d: 1/r_sp_per; SINK_C_d - SINK_d: 1/(r_hs + r_hs2)
      ambient: sink unit 1/(r_sink + r_convec s_sink^2 / A), SINK_C_d
          1/(r_hs_c_per + r_amb_c_per), SINK_d 1/(r_hs_per + r_amb_per)

    B = Laplacian of the links + diag(ambient conductances); the ambient
    enters as power amb * g_amb at the sink and package nodes
    (set_internal_power_block :382-389).  Steady: B T = P (a dense solve, the
    reference factorises B once, RCutil.c:42-183).  Transient: dT/dt =
    A^-1 P - (A^-1 B) T with the reference's adaptive RK4 (the same
    controller as the grid model, RCutil.c:440-640, compute_temp_block
    :449-486); A holds the node capacitances (populate_C_model_block
    :303-376).

    Pinned in tests/test_block_model.py against the reference library built
    from its sources: B, A and the package R/C to 1e-13 relative, steady and
    transient temperatures to 1e-9 K, plus the invariants (B symmetric, row
    sums equal to the ambient conductances, energy balance).
    """

    def __init__(self, config: dict | None = None, flp: str | None = None, units=None):
        cfg = dict(DEFAULT_CONFIG)
        cfg.update(dict(r_convec_sec=1.0, c_convec_sec=140.4, block_omit_lateral=0))
        cfg.update(config or {})
        self.cfg = cfg
        self.units = units if units is not None else read_flp(flp)
        self.n_units = n = len(self.units)
        self.N = 4 * n + EXTRA
        self.width, self.height = total_width(self.units), total_height(self.units)
        self._assemble_B()
        self._assemble_A()

    def _assemble_B(self):
        cfg, U, n, N = self.cfg, self.units, self.n_units, self.N
        w, h = self.width, self.height
        if w > cfg["s_sink"] or h > cfg["s_sink"] or w > cfg["s_spreader"] or h > cfg["s_spreader"]:
            raise ValueError("inordinate floorplan size!")
        layers = [(cfg["k_chip"], cfg["t_chip"]), (cfg["k_interface"], cfg["t_interface"]),
                  (cfg["k_spreader"], cfg["t_spreader"]), (cfg["k_sink"], cfg["t_sink"])]
        # centre-to-edge lateral conductance of every unit per layer:
        # across half the width (gx) and half the height (gy)
        gx = np.zeros((4, n))
        gy = np.zeros((4, n))
        for l, (k, t) in enumerate(layers):
            if l == 0 and cfg["block_omit_lateral"]:
                continue
            for i, u in enumerate(U):
                gx[l, i] = 1.0 / getr(k, u.width / 2.0, u.height * t)
                gy[l, i] = 1.0 / getr(k, u.height / 2.0, u.width * t)
        self.gx, self.gy = gx, gy
        pk = package_R(cfg, w, h)
        self.pack_R = pk
        B = np.zeros((N, N))

        def link(a, b, g):
            if g == 0.0:
                return
            B[a, b] -= g
            B[b, a] -= g
            B[a, a] += g
            B[b, b] += g

        for i in range(n):
            for j in range(i + 1, n):
                kind, L = unit_neighbours(U[i], U[j])
                if kind is None or L == 0.0:
                    continue
                for l in range(4):
                    if kind == "x":
                        gi, gj = gx[l, i] / U[i].height * L, gx[l, j] / U[j].height * L
                    else:
                        gi, gj = gy[l, i] / U[i].width * L, gy[l, j] / U[j].width * L
                    if gi != 0.0 and gj != 0.0:
                        link(l * n + i, l * n + j, 1.0 / (1.0 / gi + 1.0 / gj))
        X = 4 * n
        gamb = np.zeros(N)
        # edge membership (west, east, north, south) and the edge sums
        edge = np.zeros((n, 4), dtype=bool)
        for i, u in enumerate(U):
            edge[i] = [eq(u.leftx, 0), eq(u.leftx + u.width, w), eq(u.bottomy + u.height, h), eq(u.bottomy, 0)]
        self.border = edge.astype(int)
        sides = [(SP_W, SINK_C_W, "x"), (SP_E, SINK_C_E, "x"), (SP_N, SINK_C_N, "y"), (SP_S, SINK_C_S, "y")]
        for i, u in enumerate(U):
            area = u.width * u.height
            link(i, n + i, 1.0 / getr(cfg["k_chip"], cfg["t_chip"], area))
            link(n + i, 2 * n + i, 1.0 / getr(cfg["k_interface"], cfg["t_interface"], area))
            link(2 * n + i, 3 * n + i, 1.0 / getr(cfg["k_spreader"], cfg["t_spreader"], area))
            gamb[3 * n + i] = 1.0 / (getr(cfg["k_sink"], cfg["t_sink"], area)
                                     + cfg["r_convec"] * cfg["s_sink"] ** 2 / area)
            for e, (sp_node, hs_node, d) in enumerate(sides):
                if not edge[i, e]:
                    continue
                gl = gx if d == "x" else gy
                r_sp1 = pk["r_sp1_" + d]
                r_hs1 = pk["r_hs1_" + d]
                g_sp_edge = gl[2, edge[:, e]].sum()
                g_hs_edge = gl[3, edge[:, e]].sum()
                link(2 * n + i, X + sp_node, 1.0 / (1.0 / gl[2, i] + r_sp1 * g_sp_edge / gl[2, i]))
                link(3 * n + i, X + hs_node, 1.0 / (1.0 / gl[3, i] + r_hs1 * g_hs_edge / gl[3, i]))
        for sp_node, hsc, hso, d in ((SP_N, SINK_C_N, SINK_N, "y"), (SP_S, SINK_C_S, SINK_S, "y"),
                                     (SP_E, SINK_C_E, SINK_E, "x"), (SP_W, SINK_C_W, SINK_W, "x")):
            link(X + sp_node, X + hsc, 1.0 / pk["r_sp_per_" + d])
            link(X + hsc, X + hso, 1.0 / (pk["r_hs"] + pk["r_hs2_" + d]))
            gamb[X + hsc] = 1.0 / (pk["r_hs_c_per_" + d] + pk["r_amb_c_per_" + d])
            gamb[X + hso] = 1.0 / (pk["r_hs_per"] + pk["r_amb_per"])
        B += np.diag(gamb)
        self.B, self.gamb = B, gamb

    def _assemble_A(self):
        cfg, U, n, N = self.cfg, self.units, self.n_units, self.N
        pk = package_C(cfg, self.width, self.height)
        self.pack_C = pk
        a = np.zeros(N)
        for i, u in enumerate(U):
            area = u.width * u.height
            a[i] = getcap(cfg["p_chip"], cfg["t_chip"], area)
            a[n + i] = getcap(cfg["p_interface"], cfg["t_interface"], area)
            a[2 * n + i] = getcap(cfg["p_spreader"], cfg["t_spreader"], area)
            a[3 * n + i] = (getcap(cfg["p_sink"], cfg["t_sink"], area)
                            + C_FACTOR * cfg["c_convec"] / cfg["s_sink"] ** 2 * area)
        a[4 * n:] = package_node_caps(pk, False)
        self.A = a
        self.Cm = (1.0 / a)[:, None] * self.B

    def node_power(self, unit_power):
        """Node power vector: the units' power on the silicon nodes, the
        ambient's injection amb * g_amb on sink and package nodes."""
        P = self.cfg["ambient"] * self.gamb
        P[:self.n_units] = unit_power[:self.n_units]
        return P

    def steady(self, unit_power):
        return np.linalg.solve(self.B, self.node_power(unit_power))

    def slope(self, T, P):
        """dT/dt = A^-1 P - C T  (slope_fn_block, temperature_block.c:416-446)."""
        return P / self.A - self.Cm @ T

    # the adaptive RK4 and the interval loop are the grid model's, over this slope
    rk4_core = OracleModel.rk4_core
    rk4 = OracleModel.rk4
    advance = OracleModel.advance

    def transient(self, unit_power, T, dt):
        """compute_temp_block: returns (T after dt, RK4 calls)."""
        return self.advance(np.asarray(T, dtype=float), self.node_power(unit_power), dt)


def package_R(cfg, width, height):
    """populate_package_R (temperature.c:458-523)."""
    s_sp, t_sp, s_hs, t_hs = cfg["s_spreader"], cfg["t_spreader"], cfg["s_sink"], cfg["t_sink"]
    k_hs, k_sp, r_conv = cfg["k_sink"], cfg["k_spreader"], cfg["r_convec"]
    s_sub, t_sub, s_sol, t_sol = cfg["s_sub"], cfg["t_sub"], cfg["s_solder"], cfg["t_solder"]
    s_pcb, t_pcb, r_sec = cfg["s_pcb"], cfg["t_pcb"], cfg["r_convec_sec"]
    w, h = width, height
    p = {}
    p["r_sp1_x"] = getr(k_sp, (s_sp - w) / 4.0, (s_sp + 3 * h) / 4.0 * t_sp)
    p["r_sp1_y"] = getr(k_sp, (s_sp - h) / 4.0, (s_sp + 3 * w) / 4.0 * t_sp)
    p["r_hs1_x"] = getr(k_hs, (s_sp - w) / 4.0, (s_sp + 3 * h) / 4.0 * t_hs)
    p["r_hs1_y"] = getr(k_hs, (s_sp - h) / 4.0, (s_sp + 3 * w) / 4.0 * t_hs)
    p["r_hs2_x"] = getr(k_hs, (s_sp - w) / 4.0, (3 * s_sp + h) / 4.0 * t_hs)
    p["r_hs2_y"] = getr(k_hs, (s_sp - h) / 4.0, (3 * s_sp + w) / 4.0 * t_hs)
    p["r_hs"] = getr(k_hs, (s_hs - s_sp) / 4.0, (s_hs + 3 * s_sp) / 4.0 * t_hs)
    p["r_sp_per_x"] = getr(k_sp, t_sp, (s_sp + h) * (s_sp - w) / 4.0)
    p["r_sp_per_y"] = getr(k_sp, t_sp, (s_sp + w) * (s_sp - h) / 4.0)
    p["r_hs_c_per_x"] = getr(k_hs, t_hs, (s_sp + h) * (s_sp - w) / 4.0)
    p["r_hs_c_per_y"] = getr(k_hs, t_hs, (s_sp + w) * (s_sp - h) / 4.0)
    p["r_hs_per"] = getr(k_hs, t_hs, (s_hs * s_hs - s_sp * s_sp) / 4.0)
    p["r_amb_c_per_x"] = r_conv * (s_hs * s_hs) / ((s_sp + h) * (s_sp - w) / 4.0)
    p["r_amb_c_per_y"] = r_conv * (s_hs * s_hs) / ((s_sp + w) * (s_sp - h) / 4.0)
    p["r_amb_per"] = r_conv * (s_hs * s_hs) / ((s_hs * s_hs - s_sp * s_sp) / 4.0)
    p["r_sub1_x"] = getr(K_SUB, (s_sub - w) / 4.0, (s_sub + 3 * h) / 4.0 * t_sub)
    p["r_sub1_y"] = getr(K_SUB, (s_sub - h) / 4.0, (s_sub + 3 * w) / 4.0 * t_sub)
    p["r_solder1_x"] = getr(K_SOLDER, (s_sol - w) / 4.0, (s_sol + 3 * h) / 4.0 * t_sol)
    p["r_solder1_y"] = getr(K_SOLDER, (s_sol - h) / 4.0, (s_sol + 3 * w) / 4.0 * t_sol)
    p["r_pcb1_x"] = getr(K_PCB, (s_sol - w) / 4.0, (s_sol + 3 * h) / 4.0 * t_pcb)
    p["r_pcb1_y"] = getr(K_PCB, (s_sol - h) / 4.0, (s_sol + 3 * w) / 4.0 * t_pcb)
    p["r_pcb2_x"] = getr(K_PCB, (s_sol - w) / 4.0, (3 * s_sol + h) / 4.0 * t_pcb)
    p["r_pcb2_y"] = getr(K_PCB, (s_sol - h) / 4.0, (3 * s_sol + w) / 4.0 * t_pcb)
    p["r_pcb"] = getr(K_PCB, (s_pcb - s_sol) / 4.0, (s_pcb + 3 * s_sol) / 4.0 * t_pcb)
    p["r_sub_per_x"] = getr(K_SUB, t_sub, (s_sub + h) * (s_sub - w) / 4.0)
    p["r_sub_per_y"] = getr(K_SUB, t_sub, (s_sub + w) * (s_sub - h) / 4.0)
    p["r_solder_per_x"] = getr(K_SOLDER, t_sol, (s_sol + h) * (s_sol - w) / 4.0)
    p["r_solder_per_y"] = getr(K_SOLDER, t_sol, (s_sol + w) * (s_sol - h) / 4.0)
    p["r_pcb_c_per_x"] = getr(K_PCB, t_pcb, (s_sol + h) * (s_sol - w) / 4.0)
    p["r_pcb_c_per_y"] = getr(K_PCB, t_pcb, (s_sol + w) * (s_sol - h) / 4.0)
    p["r_pcb_per"] = getr(K_PCB, t_pcb, (s_pcb * s_pcb - s_sol * s_sol) / 4.0)
    p["r_amb_sec_c_per_x"] = r_sec * (s_pcb * s_pcb) / ((s_sol + h) * (s_sol - w) / 4.0)
    p["r_amb_sec_c_per_y"] = r_sec * (s_pcb * s_pcb) / ((s_sol + w) * (s_sol - h) / 4.0)
    p["r_amb_sec_per"] = r_sec * (s_pcb * s_pcb) / ((s_pcb * s_pcb - s_sol * s_sol) / 4.0)
    return p


def package_C(cfg, width, height):
    """populate_package_C (temperature.c:525-569)."""
    s_sp, t_sp, s_hs, t_hs = cfg["s_spreader"], cfg["t_spreader"], cfg["s_sink"], cfg["t_sink"]
    c_conv, p_hs, p_sp = cfg["c_convec"], cfg["p_sink"], cfg["p_spreader"]
    s_sub, t_sub, s_sol, t_sol = cfg["s_sub"], cfg["t_sub"], cfg["s_solder"], cfg["t_solder"]
    s_pcb, t_pcb, c_sec = cfg["s_pcb"], cfg["t_pcb"], cfg["c_convec_sec"]
    w, h = width, height
    p = {}
    p["c_sp_per_x"] = getcap(p_sp, t_sp, (s_sp + h) * (s_sp - w) / 4.0)
    p["c_sp_per_y"] = getcap(p_sp, t_sp, (s_sp + w) * (s_sp - h) / 4.0)
    p["c_hs_c_per_x"] = getcap(p_hs, t_hs, (s_sp + h) * (s_sp - w) / 4.0)
    p["c_hs_c_per_y"] = getcap(p_hs, t_hs, (s_sp + w) * (s_sp - h) / 4.0)
    p["c_hs_per"] = getcap(p_hs, t_hs, (s_hs * s_hs - s_sp * s_sp) / 4.0)
    p["c_amb_c_per_x"] = C_FACTOR * c_conv / (s_hs * s_hs) * ((s_sp + h) * (s_sp - w) / 4.0)
    p["c_amb_c_per_y"] = C_FACTOR * c_conv / (s_hs * s_hs) * ((s_sp + w) * (s_sp - h) / 4.0)
    p["c_amb_per"] = C_FACTOR * c_conv / (s_hs * s_hs) * ((s_hs * s_hs - s_sp * s_sp) / 4.0)
    p["c_sub_per_x"] = getcap(SPEC_HEAT_SUB, t_sub, (s_sub + h) * (s_sub - w) / 4.0)
    p["c_sub_per_y"] = getcap(SPEC_HEAT_SUB, t_sub, (s_sub + w) * (s_sub - h) / 4.0)
    p["c_solder_per_x"] = getcap(SPEC_HEAT_SOLDER, t_sol, (s_sol + h) * (s_sol - w) / 4.0)
    p["c_solder_per_y"] = getcap(SPEC_HEAT_SOLDER, t_sol, (s_sol + w) * (s_sol - h) / 4.0)
    p["c_pcb_c_per_x"] = getcap(SPEC_HEAT_PCB, t_pcb, (s_sol + h) * (s_sol - w) / 4.0)
    p["c_pcb_c_per_y"] = getcap(SPEC_HEAT_PCB, t_pcb, (s_sol + w) * (s_sol - h) / 4.0)
    p["c_pcb_per"] = getcap(SPEC_HEAT_PCB, t_pcb, (s_pcb * s_pcb - s_sol * s_sol) / 4.0)
    p["c_amb_sec_c_per_x"] = C_FACTOR * c_sec / (s_pcb * s_pcb) * ((s_sol + h) * (s_sol - w) / 4.0)
    p["c_amb_sec_c_per_y"] = C_FACTOR * c_sec / (s_pcb * s_pcb) * ((s_sol + w) * (s_sol - h) / 4.0)
    p["c_amb_sec_per"] = C_FACTOR * c_sec / (s_pcb * s_pcb) * ((s_pcb * s_pcb - s_sol * s_sol) / 4.0)
    return p


def package_node_caps(pk, secondary):
    """Capacitance of each package node, as divided out in slope_fn_pack
    (temperature_grid.c:2884-3091)."""
    c = np.zeros(EXTRA + (EXTRA_SEC if secondary else 0))
    c[SINK_N] = c[SINK_S] = c[SINK_W] = c[SINK_E] = pk["c_hs_per"] + pk["c_amb_per"]
    c[SINK_C_N] = c[SINK_C_S] = pk["c_hs_c_per_y"] + pk["c_amb_c_per_y"]
    c[SINK_C_W] = c[SINK_C_E] = pk["c_hs_c_per_x"] + pk["c_amb_c_per_x"]
    c[SP_N] = c[SP_S] = pk["c_sp_per_y"]
    c[SP_W] = c[SP_E] = pk["c_sp_per_x"]
    if secondary:
        c[PCB_N] = c[PCB_S] = c[PCB_W] = c[PCB_E] = pk["c_pcb_per"] + pk["c_amb_sec_per"]
        c[PCB_C_N] = c[PCB_C_S] = pk["c_pcb_c_per_y"] + pk["c_amb_sec_c_per_y"]
        c[PCB_C_W] = c[PCB_C_E] = pk["c_pcb_c_per_x"] + pk["c_amb_sec_c_per_x"]
        c[SOLDER_N] = c[SOLDER_S] = pk["c_solder_per_y"]
        c[SOLDER_W] = c[SOLDER_E] = pk["c_solder_per_x"]
        c[SUB_N] = c[SUB_S] = pk["c_sub_per_y"]
        c[SUB_W] = c[SUB_E] = pk["c_sub_per_x"]
    return c
