"""ctypes mirror of HotSpot's C interface for the grid thermal model.

The struct layouts follow the reference *implementation* headers, not the
stale ``hotspot-iface.h`` copy (which lacks the four per-unit fields of
``flp.h`` and orders ``grid_transient_file`` differently):

* ``thermal_config_t``      temperature.h:141-221
* ``package_RC_t``          temperature.h:237-334
* ``RC_model_t``            temperature.h:347-357
* ``unit_t`` / ``flp_t``    flp.h:378-400
* ``blist_t`` / ``glist_t`` / ``layer_t`` / ``grid_model_vector_t`` /
  ``grid_model_t``          temperature_grid.h:63-191 (SUPERLU=0 build)
* ``microchannel_config_t`` microchannel.h:29-90
* ``materials_list_t``      materials.h:10-28

The product library (``libhotspot_amd.so``) exports the same entry points
with the same layouts, so :class:`HotSpotLib` drives either library; a caller
compiled against the reference headers links against either one.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

STR_SIZE = 512
EXTRA = 12          # temperature.h:38
EXTRA_SEC = 16      # temperature.h:56
NULLFILE = b"(null)"

c_double_p = C.POINTER(C.c_double)


def _chars(n=STR_SIZE):
    return C.c_char * n


class ThermalConfig(C.Structure):
    """thermal_config_t (temperature.h:141-221)."""
    _fields_ = [
        ("t_chip", C.c_double), ("k_chip", C.c_double), ("p_chip", C.c_double),
        ("thermal_threshold", C.c_double),
        ("c_convec", C.c_double), ("r_convec", C.c_double), ("s_sink", C.c_double),
        ("t_sink", C.c_double), ("k_sink", C.c_double), ("p_sink", C.c_double),
        ("s_spreader", C.c_double), ("t_spreader", C.c_double),
        ("k_spreader", C.c_double), ("p_spreader", C.c_double),
        ("t_interface", C.c_double), ("k_interface", C.c_double),
        ("p_interface", C.c_double),
        ("model_secondary", C.c_int), ("r_convec_sec", C.c_double),
        ("c_convec_sec", C.c_double), ("n_metal", C.c_int), ("t_metal", C.c_double),
        ("t_c4", C.c_double), ("s_c4", C.c_double), ("n_c4", C.c_int),
        ("s_sub", C.c_double), ("t_sub", C.c_double), ("s_solder", C.c_double),
        ("t_solder", C.c_double), ("s_pcb", C.c_double), ("t_pcb", C.c_double),
        ("ambient", C.c_double), ("init_file", _chars()), ("init_temp", C.c_double),
        ("steady_file", _chars()), ("sampling_intvl", C.c_double),
        ("base_proc_freq", C.c_double), ("dtm_used", C.c_int),
        ("model_type", _chars()),
        ("leakage_used", C.c_int), ("leakage_mode", C.c_int),
        ("package_model_used", C.c_int), ("package_config_file", _chars()),
        ("block_omit_lateral", C.c_int),
        ("grid_rows", C.c_int), ("grid_cols", C.c_int),
        ("grid_layer_file", _chars()), ("grid_steady_file", _chars()),
        ("grid_map_mode", _chars()), ("grid_transient_file", _chars()),
        ("detailed_3D_used", C.c_int),
    ]


PACKAGE_FIELDS = [
    "r_sp1_x", "r_sp1_y", "r_hs1_x", "r_hs1_y", "r_hs2_x", "r_hs2_y", "r_hs",
    "r_sp_per_x", "r_sp_per_y", "r_hs_c_per_x", "r_hs_c_per_y", "r_hs_per",
    "c_sp_per_x", "c_sp_per_y", "c_hs_c_per_x", "c_hs_c_per_y", "c_hs_per",
    "r_amb_c_per_x", "c_amb_c_per_x", "r_amb_c_per_y", "c_amb_c_per_y",
    "r_amb_per", "c_amb_per",
    "r_sub1_x", "r_sub1_y", "r_solder1_x", "r_solder1_y", "r_pcb1_x", "r_pcb1_y",
    "r_pcb2_x", "r_pcb2_y", "r_pcb",
    "r_sub_per_x", "r_sub_per_y", "r_solder_per_x", "r_solder_per_y",
    "r_pcb_c_per_x", "r_pcb_c_per_y", "r_pcb_per",
    "c_sub_per_x", "c_sub_per_y", "c_solder_per_x", "c_solder_per_y",
    "c_pcb_c_per_x", "c_pcb_c_per_y", "c_pcb_per",
    "r_amb_sec_c_per_x", "c_amb_sec_c_per_x", "r_amb_sec_c_per_y",
    "c_amb_sec_c_per_y", "r_amb_sec_per", "c_amb_sec_per",
]


class PackageRC(C.Structure):
    """package_RC_t (temperature.h:237-334): 52 doubles."""
    _fields_ = [(n, C.c_double) for n in PACKAGE_FIELDS]


class Unit(C.Structure):
    """unit_t (flp.h:378-391), including the per-unit detailed-3D fields."""
    _fields_ = [
        ("name", _chars()), ("width", C.c_double), ("height", C.c_double),
        ("leftx", C.c_double), ("bottomy", C.c_double),
        ("specificheat", C.c_double), ("resistivity", C.c_double),
        ("hasRes", C.c_int), ("hasSh", C.c_int),
    ]


class Flp(C.Structure):
    """flp_t (flp.h:394-400)."""
    _fields_ = [("units", C.POINTER(Unit)), ("n_units", C.c_int),
                ("wire_density", C.POINTER(c_double_p))]


class MicrochannelConfig(C.Structure):
    """microchannel_config_t (microchannel.h:29-90)."""
    _fields_ = [
        ("cell_width", C.c_double), ("cell_height", C.c_double),
        ("cell_thickness", C.c_double), ("pumping_pressure", C.c_double),
        ("pump_internal_res", C.c_double), ("inlet_temperature", C.c_double),
        ("coolant_capac", C.c_double), ("coolant_res", C.c_double),
        ("coolant_visc", C.c_double), ("wall_capac", C.c_double),
        ("wall_res", C.c_double), ("htc", C.c_double),
        ("network_file", _chars()), ("floorplan_file", _chars(STR_SIZE + 3)),
        ("num_rows", C.c_int), ("num_columns", C.c_int), ("n_fluid_cells", C.c_int),
        ("cell_types", C.POINTER(C.POINTER(C.c_int))),
        ("mapping", C.POINTER(C.POINTER(C.c_int))),
        ("A", C.POINTER(c_double_p)), ("b", c_double_p), ("nnz", C.c_int),
    ]


class Material(C.Structure):
    _fields_ = [("material_type", C.c_int), ("thermal_conductivity", C.c_double),
                ("volumetric_heat_capacity", C.c_double),
                ("dynamic_viscosity", C.c_double)]


class MaterialsList(C.Structure):
    """materials_list_t (materials.h:18-28)."""
    _fields_ = [("size", C.c_int), ("names", C.POINTER(C.c_char_p)),
                ("material_properties", C.POINTER(Material))]


class Blist(C.Structure):
    pass


Blist._fields_ = [
    ("idx", C.c_int), ("occupancy", C.c_double), ("next", C.POINTER(Blist)),
    ("lock", C.c_int), ("hasRes", C.c_int), ("hasCap", C.c_int),
    ("rx", C.c_double), ("ry", C.c_double), ("rz", C.c_double),
    ("capacitance", C.c_double),
]


class Glist(C.Structure):
    _fields_ = [("i1", C.c_int), ("i2", C.c_int), ("j1", C.c_int), ("j2", C.c_int)]


class Layer(C.Structure):
    """layer_t (temperature_grid.h:101-126)."""
    _fields_ = [
        ("flp", C.POINTER(Flp)), ("no", C.c_int), ("has_lateral", C.c_int),
        ("has_power", C.c_int), ("k", C.c_double), ("thickness", C.c_double),
        ("sp", C.c_double), ("is_microchannel", C.c_int),
        ("microchannel_config", C.POINTER(MicrochannelConfig)),
        ("rx", C.c_double), ("ry", C.c_double), ("rz", C.c_double), ("c", C.c_double),
        ("b2gmap", C.POINTER(C.POINTER(C.POINTER(Blist)))),
        ("g2bmap", C.POINTER(Glist)),
    ]


class GridVector(C.Structure):
    """grid_model_vector_t (temperature_grid.h:129-135)."""
    _fields_ = [("cuboid", C.POINTER(C.POINTER(c_double_p))), ("extra", c_double_p)]


class GridModel(C.Structure):
    """grid_model_t (temperature_grid.h:138-191), SUPERLU=0 layout."""
    _fields_ = [
        ("config", ThermalConfig), ("layers", C.POINTER(Layer)), ("n_layers", C.c_int),
        ("rows", C.c_int), ("cols", C.c_int), ("width", C.c_double),
        ("height", C.c_double), ("pack", PackageRC), ("total_n_blocks", C.c_int),
        ("map_mode", C.c_int), ("r_ready", C.c_int), ("c_ready", C.c_int),
        ("has_lcf", C.c_int), ("last_steady", C.POINTER(GridVector)),
        ("last_trans", C.POINTER(GridVector)), ("last_temp", c_double_p),
        ("base_n_units", C.c_int), ("use_microchannels", C.c_int),
        ("default_microchannel_config", C.POINTER(MicrochannelConfig)),
    ]


class BlockModel(C.Structure):
    """block_model_t (temperature_block.h:16-67)."""
    _fields_ = [
        ("flp", C.POINTER(Flp)), ("config", ThermalConfig),
        ("b", C.POINTER(c_double_p)), ("lu", C.POINTER(c_double_p)), ("p", C.POINTER(C.c_int)),
        ("a", c_double_p), ("inva", c_double_p), ("c", C.POINTER(c_double_p)),
        ("pack", PackageRC),
        ("gx", c_double_p), ("gy", c_double_p), ("gx_int", c_double_p), ("gy_int", c_double_p),
        ("gx_sp", c_double_p), ("gy_sp", c_double_p), ("gx_hs", c_double_p), ("gy_hs", c_double_p),
        ("g_amb", c_double_p), ("t_vector", c_double_p),
        ("len", C.POINTER(c_double_p)), ("g", C.POINTER(c_double_p)),
        ("border", C.POINTER(C.POINTER(C.c_int))),
        ("n_nodes", C.c_int), ("n_units", C.c_int), ("base_n_units", C.c_int),
        ("r_ready", C.c_int), ("c_ready", C.c_int),
    ]


class RCModel(C.Structure):
    """RC_model_t (temperature.h:347-357); the union holds a grid_model_t* or,
    for type BLOCK_MODEL (0), a block_model_t* (see ``block_of``)."""
    _fields_ = [("grid", C.POINTER(GridModel)), ("type", C.c_int),
                ("config", C.POINTER(ThermalConfig))]


BLOCK_MODEL, GRID_MODEL = 0, 1


def block_of(model) -> BlockModel:
    """The block_model_t of an RC_model_t of type BLOCK_MODEL."""
    return C.cast(model.contents.grid, C.POINTER(BlockModel)).contents


def block_matrix(ptr, n: int) -> np.ndarray:
    """Copy of an n x n dmatrix (row pointers into one contiguous block)."""
    return np.ctypeslib.as_array(ptr[0], shape=(n * n,)).reshape(n, n).copy()


RCModelP = C.POINTER(RCModel)
FlpP = C.POINTER(Flp)


class PackageConfig(C.Structure):
    """package_config_t (package.h:48-75)"""
    _fields_ = [("natural_convec", C.c_int), ("flow_type", C.c_int), ("sink_type", C.c_int),
                ("fin_height", C.c_double), ("fin_width", C.c_double), ("channel_width", C.c_double),
                ("pin_height", C.c_double), ("pin_diam", C.c_double), ("pin_dist", C.c_double),
                ("fan_radius", C.c_double), ("motor_radius", C.c_double), ("rpm", C.c_int)]


class Convection(C.Structure):
    """convection_t (package.h:78-104)"""
    _fields_ = [(n, C.c_double) for n in ("n_fin", "sur_area_fin", "n_pin", "sur_area_pin", "reynolds",
                                          "nusselt", "h_coeff", "v", "r_th")]


class StrPair(C.Structure):
    """str_pair (util.h:92-96)"""
    _fields_ = [("name", C.c_char * STR_SIZE), ("value", C.c_char * STR_SIZE)]


def str_table(pairs, size: int = 512):
    """A str_pair array holding (name, value) pairs."""
    t = (StrPair * size)()
    for i, (k, v) in enumerate(pairs):
        t[i].name = k.encode()
        t[i].value = v.encode()
    return t


class HotSpotLib:
    """Loads a library exporting HotSpot's C API and binds its signatures.

    The same object type drives the reference library (oracle/_ref) and the
    product library; nothing here depends on which one is loaded.
    """

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = C.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L = self.lib
        L.default_thermal_config.restype = ThermalConfig
        L.default_thermal_config.argtypes = []
        L.default_microchannel_config.restype = MicrochannelConfig
        L.default_microchannel_config.argtypes = []
        L.default_materials.argtypes = [C.POINTER(MaterialsList)]
        L.default_materials.restype = None
        L.read_flp.restype = FlpP
        L.read_flp.argtypes = [C.c_char_p, C.c_int, C.c_int]
        L.free_flp.restype = None
        L.free_flp.argtypes = [FlpP, C.c_int, C.c_int]
        L.alloc_RC_model.restype = RCModelP
        L.alloc_RC_model.argtypes = [C.POINTER(ThermalConfig), FlpP,
                                     C.POINTER(MicrochannelConfig),
                                     C.POINTER(MaterialsList), C.c_int, C.c_int]
        L.delete_RC_model.restype = None
        L.delete_RC_model.argtypes = [RCModelP]
        for fn in ("populate_R_model", "populate_C_model"):
            getattr(L, fn).restype = None
            getattr(L, fn).argtypes = [RCModelP, FlpP]
        L.hotspot_vector.restype = c_double_p
        L.hotspot_vector.argtypes = [RCModelP]
        L.free_dvector.restype = None
        L.free_dvector.argtypes = [c_double_p]
        L.set_temp.restype = None
        L.set_temp.argtypes = [RCModelP, c_double_p, C.c_double]
        L.copy_temp.restype = None
        L.copy_temp.argtypes = [RCModelP, c_double_p, c_double_p]
        L.steady_state_temp.restype = None
        L.steady_state_temp.argtypes = [RCModelP, c_double_p, c_double_p]
        L.compute_temp.restype = None
        L.compute_temp.argtypes = [RCModelP, c_double_p, c_double_p, C.c_double]
        L.dump_temp.restype = None
        L.dump_temp.argtypes = [RCModelP, c_double_p, C.c_char_p]
        L.read_temp.restype = None
        L.read_temp.argtypes = [RCModelP, c_double_p, C.c_char_p, C.c_int]
        L.read_power.restype = None
        L.read_power.argtypes = [RCModelP, c_double_p, C.c_char_p]
        L.dump_power.restype = None
        L.dump_power.argtypes = [RCModelP, c_double_p, C.c_char_p]
        L.find_max_temp.restype = C.c_double
        L.find_max_temp.argtypes = [RCModelP, c_double_p]
        L.find_avg_temp.restype = C.c_double
        L.find_avg_temp.argtypes = [RCModelP, c_double_p]
        L.calc_sink_temp.restype = C.c_double
        L.calc_sink_temp.argtypes = [RCModelP, c_double_p]
        L.get_blk_index.restype = C.c_int
        L.get_blk_index.argtypes = [FlpP, C.c_char_p]
        L.dump_steady_temp_grid.restype = None
        L.dump_steady_temp_grid.argtypes = [C.POINTER(GridModel), C.c_char_p]
        # configuration tables (util.c:322-439, temperature.c:115-402, microchannel.c:49-156)
        SP = C.POINTER(StrPair)
        L.read_str_pairs.argtypes = [SP, C.c_int, C.c_char_p]
        L.read_str_pairs.restype = C.c_int
        L.parse_cmdline.argtypes = [SP, C.c_int, C.c_int, C.POINTER(C.c_char_p)]
        L.parse_cmdline.restype = C.c_int
        L.str_pairs_remove_duplicates.argtypes = [SP, C.c_int]
        L.str_pairs_remove_duplicates.restype = C.c_int
        L.get_str_index.argtypes = [SP, C.c_int, C.c_char_p]
        L.get_str_index.restype = C.c_int
        L.dump_str_pairs.argtypes = [SP, C.c_int, C.c_char_p, C.c_char_p]
        L.dump_str_pairs.restype = None
        L.thermal_config_add_from_strs.argtypes = [C.POINTER(ThermalConfig), C.POINTER(MaterialsList), SP,
                                                   C.c_int]
        L.thermal_config_add_from_strs.restype = None
        L.thermal_config_to_strs.argtypes = [C.POINTER(ThermalConfig), SP, C.c_int]
        L.thermal_config_to_strs.restype = C.c_int
        L.microchannel_config_add_from_strs.argtypes = [C.POINTER(MicrochannelConfig), C.POINTER(MaterialsList),
                                                        SP, C.c_int]
        L.microchannel_config_add_from_strs.restype = None
        L.microchannel_config_to_strs.argtypes = [C.POINTER(MicrochannelConfig), SP, C.c_int]
        L.microchannel_config_to_strs.restype = C.c_int
        L.materials_add_from_file.argtypes = [C.POINTER(MaterialsList), C.c_char_p]
        L.materials_add_from_file.restype = None

    # ---- thin wrappers -------------------------------------------------
    def default_config(self, **overrides) -> ThermalConfig:
        cfg = self.lib.default_thermal_config()
        apply_overrides(cfg, overrides)
        return cfg

    def n_vector(self, model: RCModelP) -> int:
        if model.contents.type == BLOCK_MODEL:
            return block_of(model).n_nodes
        g = model.contents.grid.contents
        extra = EXTRA + (EXTRA_SEC if g.config.model_secondary else 0)
        return g.total_n_blocks + extra

    def vector(self, model: RCModelP):
        """hotspot_vector() plus a numpy view over it."""
        p = self.lib.hotspot_vector(model)
        return p, np.ctypeslib.as_array(p, shape=(self.n_vector(model),))


def apply_overrides(cfg: ThermalConfig, overrides: dict) -> None:
    names = {f[0]: f[1] for f in ThermalConfig._fields_}
    for k, v in overrides.items():
        if k not in names:
            raise KeyError(f"unknown thermal_config_t field {k!r}")
        if isinstance(v, str):
            v = v.encode()
        setattr(cfg, k, v)


def grid_vector_array(g: GridModel, gv_ptr) -> np.ndarray:
    """Copy of a grid_model_vector_t as a flat [nl*nr*nc + extras] array."""
    extra = EXTRA + (EXTRA_SEC if g.config.model_secondary else 0)
    n = g.n_layers * g.rows * g.cols + extra
    base = gv_ptr.contents.cuboid[0][0]
    return np.ctypeslib.as_array(base, shape=(n,)).copy()
