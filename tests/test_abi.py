"""The C ABI: every function declared in include/hotspot_amd.h is exported by
the built library; the struct layouts of the header equal the ctypes mirror
the tests use and, where the reference sources are present, the layouts the
reference's own headers give (a probe compiled once against each header set,
every field offset and size compared); a C program compiled against the
header runs the host-side API."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from hotspot_amd import iface

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hotspot_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("sizeof",)))


def test_every_declared_function_is_exported(prod):
    names = declared_functions()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(prod.lib, n)]
    assert not missing, missing


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "hotspot_amd.h"
#define P(T, F) printf(#T "." #F " %zu\n", offsetof(T, F));
#define S(T) printf(#T " %zu\n", sizeof(T));
int main(void) {
  S(thermal_config_t) S(unit_t) S(flp_t) S(package_RC_t) S(layer_t) S(grid_model_t)
  S(RC_model_t) S(microchannel_config_t) S(blist_t) S(glist_t) S(package_config_t) S(convection_t)
  P(package_config_t, rpm) P(convection_t, r_th)
  P(thermal_config_t, model_secondary) P(thermal_config_t, ambient) P(thermal_config_t, grid_rows)
  P(thermal_config_t, grid_map_mode) P(thermal_config_t, grid_transient_file)
  P(thermal_config_t, detailed_3D_used) P(layer_t, flp) P(layer_t, has_power) P(layer_t, rx)
  P(layer_t, c) P(layer_t, g2bmap) P(grid_model_t, layers) P(grid_model_t, n_layers)
  P(grid_model_t, pack) P(grid_model_t, total_n_blocks) P(grid_model_t, has_lcf)
  P(grid_model_t, last_temp) P(grid_model_t, default_microchannel_config)
  P(RC_model_t, type) P(RC_model_t, config) P(unit_t, hasSh) P(microchannel_config_t, nnz)
  return 0;
}
"""

CTYPES = {"thermal_config_t": iface.ThermalConfig, "unit_t": iface.Unit, "flp_t": iface.Flp,
          "package_RC_t": iface.PackageRC, "layer_t": iface.Layer, "grid_model_t": iface.GridModel,
          "RC_model_t": iface.RCModel, "microchannel_config_t": iface.MicrochannelConfig,
          "blist_t": iface.Blist, "glist_t": iface.Glist, "package_config_t": iface.PackageConfig,
          "convection_t": iface.Convection}


def test_struct_layouts_match_reference_mirror():
    tmp = tempfile.mkdtemp()
    src = os.path.join(tmp, "probe.c")
    exe = os.path.join(tmp, "probe")
    open(src, "w").write(PROBE)
    subprocess.check_call(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), src, "-o", exe])
    for line in subprocess.check_output([exe]).decode().split("\n"):
        if not line:
            continue
        key, val = line.split()
        if "." in key:
            t, f = key.split(".")
            assert getattr(CTYPES[t], f).offset == int(val), key
        else:
            assert C.sizeof(CTYPES[key]) == int(val), key


REF_DIR = "/root/reference"
# typedef name -> ctypes mirror (which lists every field) or the field names
REF_STRUCTS = {
    "thermal_config_t": iface.ThermalConfig, "unit_t": iface.Unit, "flp_t": iface.Flp,
    "package_RC_t": iface.PackageRC, "layer_t": iface.Layer, "grid_model_t": iface.GridModel,
    "grid_model_vector_t": iface.GridVector, "RC_model_t": iface.RCModel,
    "block_model_t": iface.BlockModel, "microchannel_config_t": iface.MicrochannelConfig,
    "material_t": iface.Material, "materials_list_t": iface.MaterialsList, "blist_t": iface.Blist,
    "glist_t": iface.Glist, "str_pair": iface.StrPair,
    "package_config_t": iface.PackageConfig, "convection_t": iface.Convection,
}


def layout_probe(includes):
    lines = ["#include <stddef.h>", "#include <stdio.h>"] + [f'#include "{h}"' for h in includes]
    lines.append("int main(void) {")
    for t, spec in REF_STRUCTS.items():
        fields = [f[0] for f in spec._fields_] if hasattr(spec, "_fields_") else spec
        lines.append(f'  printf("{t} %zu\\n", sizeof({t}));')
        for f in fields:
            lines.append(f'  printf("{t}.{f} %zu\\n", offsetof({t}, {f}));')
    lines += ["  return 0;", "}"]
    return "\n".join(lines) + "\n"


def run_probe(src_text, flags, tmp):
    src = os.path.join(tmp, "probe.c")
    exe = os.path.join(tmp, "probe")
    with open(src, "w") as f:
        f.write(src_text)
    subprocess.check_call(["gcc", "-std=gnu11", "-w"] + flags + [src, "-o", exe])
    return subprocess.check_output([exe]).decode().split("\n")


@pytest.mark.skipif(not os.path.isdir(REF_DIR), reason="reference sources absent")
def test_struct_layouts_match_reference_headers(tmp_path):
    """Every field offset and struct size of include/hotspot_amd.h equals the
    one the reference's own headers give (temperature.h:141-357,
    temperature_grid.h:63-191, temperature_block.h:18-62, flp.h:122-144,
    microchannel.h:29-90, materials.h:9-27, util.h:92, package.h:48-104),
    compiled as the reference compiles them (no SuperLU, no math library)."""
    (tmp_path / "ours").mkdir()
    ours = run_probe(layout_probe(["hotspot_amd.h"]), ["-I" + os.path.join(ROOT, "include")],
                     str(tmp_path / "ours"))
    theirs = run_probe(layout_probe(["temperature.h", "temperature_grid.h", "temperature_block.h",
                                     "package.h"]),
                       ["-I" + REF_DIR, "-DSUPERLU=0", "-DMATHACCEL=0"], str(tmp_path))
    assert len(ours) > 150
    assert ours == theirs, [(a, b) for a, b in zip(ours, theirs) if a != b][:10]


DRIVER = r"""
#include <stdio.h>
#include <string.h>
#include "hotspot_amd.h"
int main(int argc, char **argv) {
  thermal_config_t cfg = default_thermal_config();
  strcpy(cfg.model_type, "grid");
  cfg.grid_rows = cfg.grid_cols = 32;
  flp_t *flp = read_flp(argv[1], 0, 0);
  RC_model_t *m = alloc_RC_model(&cfg, flp, NULL, NULL, 0, 0);
  populate_R_model(m, flp);
  populate_C_model(m, flp);
  /* the fields the reference CLI reads directly (hotspot.c:546-549, :578-587) */
  int n = 0;
  for (int i = 0; i < m->grid->n_layers; i++)
    if (m->grid->layers[i].has_power) n += m->grid->layers[i].flp->n_units;
  printf("%d %d %d %.17g\n", m->grid->n_layers, n, m->grid->has_lcf, m->grid->layers[0].rx);
  delete_RC_model(m);
  free_flp(flp, 0, 0);
  return 0;
}
"""


def test_c_caller_compiled_against_header(prod, tmp_path):
    from hotspot_amd import cases
    src = tmp_path / "drv.c"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    libdir = os.path.dirname(prod.path)
    subprocess.check_call(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                           "-L" + libdir, "-lhotspot_amd", "-Wl,-rpath," + libdir])
    fp = tmp_path / "chip.flp"
    cases.write_flp(str(fp), cases.tiled_floorplan(4, 4, 0.016, 0.016))
    out = subprocess.check_output([str(exe), str(fp)]).decode().split()
    assert out[:3] == ["4", "16", "0"]
    assert abs(float(out[3]) - 0.016 / 32 / (100.0 * 0.016 / 32 * 0.15e-3)) < 1e-12


def test_extension_struct_layouts_match_mirrors(tmp_path):
    """The structs the header adds (options, statistics, engine job and
    engine statistics records) have the field offsets and sizes of the
    ctypes mirrors that the engine wrapper, the tests and bench.py use."""
    import bench
    from hotspot_amd import engine
    from tests import helpers
    mirrors = {"hs_options_t": [engine.Options, helpers.HsOptions, bench.Opts],
               "hs_stats_t": [helpers.HsStats, bench.Stats],
               "hs_job_info_t": [engine.JobInfo], "hs_engine_stats_t": [engine.EngineStats]}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "hotspot_amd.h"', "int main(void) {"]
    for t, ms in mirrors.items():
        lines.append(f'  printf("{t} %zu\\n", sizeof({t}));')
        for f in ms[0]._fields_:
            lines.append(f'  printf("{t}.{f[0]} %zu\\n", offsetof({t}, {f[0]}));')
    lines += ["  return 0;", "}"]
    out = run_probe("\n".join(lines) + "\n", ["-I" + os.path.join(ROOT, "include")], str(tmp_path))
    n = 0
    for line in out:
        if not line:
            continue
        key, val = line.split()
        t, _, f = key.partition(".")
        for m in mirrors[t]:
            assert (getattr(m, f).offset if f else C.sizeof(m)) == int(val), (m.__name__, key)
            n += 1
    assert n > 40
