"""The configuration tables and the command-line simulator, on the CPU.

The name/value tables (read_str_pairs, parse_cmdline,
str_pairs_remove_duplicates) and the configuration parsers
(thermal_config_add_from_strs, thermal_config_to_strs,
microchannel_config_add_from_strs) are compared field by field with the
reference library compiled from its sources (oracle/_ref), on the
reference's own example configuration and materials files; the
command-line simulator's argument and error handling runs without a GPU.
The simulations themselves are in test_gpu_cli.py.
"""
import ctypes as C
import os
import subprocess

import pytest

from hotspot_amd import iface

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI_DIR = os.path.join(ROOT, "tests", "golden", "cli")
EXE = os.path.join(ROOT, "hotspot_amd", "hotspot")
CONFIGS = [os.path.join(CLI_DIR, c, "example.config") for c in ("ev6", "lcf_detailed", "ev6_3d")]
MATERIALS = os.path.join(CLI_DIR, "lcf_detailed", "example.materials")


def table_of(lib, path):
    t = (iface.StrPair * 512)()
    n = lib.lib.read_str_pairs(t, 512, path.encode())
    return t, n


def pairs(t, n):
    return [(t[i].name.decode(), t[i].value.decode()) for i in range(n)]


def fields(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        out[name] = v if not hasattr(v, "_length_") else bytes(v)
    return out


def materials(lib):
    ml = iface.MaterialsList()
    lib.lib.default_materials(C.byref(ml))
    lib.lib.materials_add_from_file(C.byref(ml), MATERIALS.encode())
    return ml


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda p: os.path.basename(os.path.dirname(p)))
def test_config_file_tables_match_reference(prod, ref, cfg):
    tp, np_ = table_of(prod, cfg)
    tr, nr = table_of(ref, cfg)
    assert np_ == nr > 40
    assert pairs(tp, np_) == pairs(tr, nr)


def test_cmdline_and_duplicates_match_reference(prod, ref):
    argv = [b"hotspot", b"-c", b"x.config", b"-model_type", b"grid", b"-GRID_ROWS", b"128", b"-grid_rows", b"32",
            b"-p", b"trace"]
    arr = (C.c_char_p * len(argv))(*argv)
    out = []
    for lib in (prod, ref):
        t = (iface.StrPair * 512)()
        n = lib.lib.parse_cmdline(t, 512, len(argv), arr)
        m = lib.lib.str_pairs_remove_duplicates(t, n)
        idx = lib.lib.get_str_index(t, m, b"grid_rows")
        out.append((n, pairs(t, m), idx))
    assert out[0] == out[1]
    assert out[0][0] == 5 and out[0][2] == 2 and dict(out[0][1])["GRID_ROWS"] == "128"


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda p: os.path.basename(os.path.dirname(p)))
def test_thermal_config_from_tables_matches_reference(prod, ref, cfg):
    """Defaults overridden by the example file plus command-line entries, with
    material names resolved through the materials file."""
    extra = [("model_type", "grid"), ("grid_rows", "96"), ("sampling_intvl", "2e-6"),
             ("material_interface", "aluminum")]
    res = []
    for lib in (prod, ref):
        # command-line entries first: they win over the file's
        t = iface.str_table(extra)
        base = C.cast(C.byref(t, len(extra) * C.sizeof(iface.StrPair)), C.POINTER(iface.StrPair))
        n = len(extra) + lib.lib.read_str_pairs(base, 512 - len(extra), cfg.encode())
        n = lib.lib.str_pairs_remove_duplicates(t, n)
        c = lib.lib.default_thermal_config()
        ml = materials(lib)
        lib.lib.thermal_config_add_from_strs(C.byref(c), C.byref(ml), t, n)
        out = (iface.StrPair * 512)()
        m = lib.lib.thermal_config_to_strs(C.byref(c), out, 512)
        res.append((fields(c), pairs(out, m)))
    assert res[0][0] == res[1][0]
    assert res[0][1] == res[1][1]
    assert len(res[0][1]) == 51
    c = res[0][0]
    assert c["k_interface"] == 237.0 and c["grid_rows"] == 96 and c["sampling_intvl"] == 2e-6


def test_microchannel_config_from_tables_matches_reference(prod, ref):
    entries = [("pumping_pressure", "12000"), ("inlet_temperature", "305.5"), ("htc", "30000"),
               ("wall_material", "silicon"), ("coolant_material", "water"), ("network_file", "net.csv")]
    res = []
    for lib in (prod, ref):
        t = iface.str_table(entries)
        c = lib.lib.default_microchannel_config()
        ml = materials(lib)
        lib.lib.microchannel_config_add_from_strs(C.byref(c), C.byref(ml), t, len(entries))
        out = (iface.StrPair * 512)()
        m = lib.lib.microchannel_config_to_strs(C.byref(c), out, 512)
        # floorplan_file is left uninitialised by the reference's default_microchannel_config
        skip = ("cell_types", "mapping", "A", "b", "floorplan_file")
        res.append(({k: v for k, v in fields(c).items() if k not in skip},
                    [(out[i].name.decode(), out[i].value.decode("latin-1")) for i in range(17) if i != 13], m))
    assert res[0] == res[1]
    c = res[0][0]
    assert c["coolant_res"] == 1.0 / 0.6069 and c["coolant_visc"] == 8.89e-4 and c["wall_res"] == 1.0 / 130.0
    assert res[0][2] == 16   # the reference counts 16 of the 17 entries it writes


def test_config_errors_are_reported(prod):
    L = prod.lib
    buf = C.create_string_buffer(512)
    for entries, msg in [([("t_chip", "-1")], "dimensions"), ([("grid_map_mode", "median")], "mapping mode"),
                         ([("model_type", "blocky")], "model type"), ([("k_chip", "abc")], "k_chip"),
                         ([("material_chip", "unobtainium")], "material_chip")]:
        c = L.default_thermal_config()
        ml = iface.MaterialsList()
        L.default_materials(C.byref(ml))
        L.thermal_config_add_from_strs(C.byref(c), C.byref(ml), iface.str_table(entries), len(entries))
        assert L.hs_last_error(buf, 512) == 1 and msg in buf.value.decode()


@pytest.fixture(scope="module")
def exe(prod):
    if not os.path.exists(EXE):
        from hotspot_amd import build
        build.build_cli(verbose=False)
    return EXE


def test_cli_usage_without_enough_arguments(exe):
    r = subprocess.run([exe, "-p", "x"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stdout


def test_cli_block_model_needs_a_floorplan(exe, tmp_path):
    """The block model is built from a floorplan file; without -f the run
    stops before any GPU use."""
    d = os.path.join(CLI_DIR, "ev6_block")
    base = [exe, "-c", os.path.join(d, "example.config"), "-p", os.path.join(d, "gcc.ptrace"),
            "-materials_file", os.path.join(d, "example.materials")]
    r = subprocess.run(base + ["-model_type", "block"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "needs a floorplan file" in r.stderr


def test_cli_dumps_its_configuration(exe, tmp_path, ref):
    """-d writes the parameters in use; the thermal part matches what the
    reference's parsers make of the same tables.  The run stops at the
    missing floorplan, after the dump and before any GPU use."""
    d = os.path.join(CLI_DIR, "ev6")
    dump = str(tmp_path / "dump.config")
    r = subprocess.run([exe, "-c", os.path.join(d, "example.config"), "-p", os.path.join(d, "gcc.ptrace"),
                        "-materials_file", os.path.join(d, "example.materials"), "-model_type", "grid",
                        "-d", dump], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "LCF or FLP" in r.stderr
    got = dict(line.rstrip("\n").split("\t", 1) for line in open(dump))
    assert got["-p"] == os.path.join(d, "gcc.ptrace") and got["-model_type"] == "grid"
    t = iface.str_table([("model_type", "grid")])
    base = C.cast(C.byref(t, C.sizeof(iface.StrPair)), C.POINTER(iface.StrPair))
    n = 1 + ref.lib.read_str_pairs(base, 511, os.path.join(d, "example.config").encode())
    n = ref.lib.str_pairs_remove_duplicates(t, n)
    c = ref.lib.default_thermal_config()
    ml2 = iface.MaterialsList()
    ref.lib.default_materials(C.byref(ml2))
    ref.lib.materials_add_from_file(C.byref(ml2), os.path.join(d, "example.materials").encode())
    ref.lib.thermal_config_add_from_strs(C.byref(c), C.byref(ml2), t, n)
    out = (iface.StrPair * 512)()
    m = ref.lib.thermal_config_to_strs(C.byref(c), out, 512)
    for k, v in pairs(out, m):
        assert got["-" + k] == v, k
