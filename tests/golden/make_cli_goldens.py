"""Writes the command-line golden files under tests/golden/cli/: the inputs of
the reference's examples 1-5 (floorplans, layer files, power traces, configs,
materials, channel networks: data files) and the outputs of the reference's
own command-line simulator on examples 1-4, run with the command lines of
the examples' run.sh.

Run in a container that holds the reference sources, after
``make -C oracle ref cli`` (which compiles the reference from its sources into
oracle/_ref/).  The GPU tests compare the command-line simulator of this
library with these outputs.
"""
from __future__ import annotations

import gzip
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/examples"
EXE = os.path.join(ROOT, "oracle", "_ref", "hotspot_ref")
OUT = os.path.join(HERE, "cli")

# case -> (example dir, input files, runs); a run is (argv, outputs to keep).
# "@init" in argv names the steady file of the case's first run as input; an
# output named "!name" marks a run that must fail, and keeps its error line.
CASES = {
    "ev6_block": ("example1", ["ev6.flp", "gcc.ptrace", "example.config", "example.materials",
                               "package.config", "@natural.config"], [
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
         "-steady_file gcc.steady -o first.ttrace", ["gcc.steady", "first.ttrace"]),
        ("-c example.config -init_file @init -f ev6.flp -p gcc.ptrace -materials_file example.materials "
         "-model_type block -o gcc.ttrace", ["gcc.ttrace"]),
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
         "-package_model_used 1 -package_config_file natural.config -steady_file natural.steady",
         ["natural.steady"]),
        # the transient starts at ambient (init_temp = ambient in example 1): a
        # natural-convection sink at ambient has no convection, B loses its
        # ground and the LU stops at a pivot below 1e-6 in the first interval
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
         "-package_model_used 1 -package_config_file natural.config -o natural.ttrace", ["!natural.err"]),
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
         "-package_model_used 1 -package_config_file package.config -steady_file forced.steady",
         ["forced.steady"]),
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
         "-leakage_used 1 -steady_file leak.steady", ["leak.steady"]),
    ]),
    "ev6": ("example2", ["ev6.flp", "gcc.ptrace", "example.config", "example.materials"], [
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type grid "
         "-steady_file gcc.steady -grid_steady_file gcc.grid.steady", ["gcc.steady", "gcc.grid.steady"]),
        ("-c example.config -init_file @init -f ev6.flp -p gcc.ptrace -materials_file example.materials "
         "-model_type grid -o gcc.ttrace", ["gcc.ttrace"]),
        # the grid model's temperature-leakage fixed point (temperature.c:708-746)
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type grid "
         "-leakage_used 1 -steady_file leak.steady -grid_steady_file leak.grid.steady",
         ["leak.steady", "leak.grid.steady"]),
    ]),
    "ev6_package": ("example2", ["ev6.flp", "gcc.ptrace", "example.config", "example.materials",
                                 "@example1/package.config", "@natural.config"], [
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type grid "
         "-package_model_used 1 -package_config_file package.config -steady_file forced.steady",
         ["forced.steady"]),
        ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type grid "
         "-package_model_used 1 -package_config_file natural.config -steady_file natural.steady "
         "-grid_steady_file natural.grid.steady", ["natural.steady", "natural.grid.steady"]),
    ]),
    "lcf_detailed": ("example3", ["example.config", "example.lcf", "example.materials", "example.ptrace",
                                  "floorplan1.flp", "floorplan2.flp"], [
        ("-c example.config -p example.ptrace -grid_layer_file example.lcf -materials_file example.materials "
         "-model_type grid -detailed_3D on -steady_file example.steady -grid_steady_file example.grid.steady",
         ["example.steady", "example.grid.steady"]),
        ("-c example.config -p example.ptrace -grid_layer_file example.lcf -materials_file example.materials "
         "-model_type grid -detailed_3D on -init_file @init -o example.ttrace "
         "-grid_transient_file example.grid.ttrace", ["example.ttrace", "example.grid.ttrace"]),
    ]),
    "ev6_3d": ("example4", ["example.config", "ev6_3D.lcf", "ev6_3D.ptrace", "ev6_3D_TIM.flp",
                            "ev6_3D_TIM_TSV.flp", "ev6_3D_cache_1.flp", "ev6_3D_cache_2.flp",
                            "ev6_3D_core_layer.flp"], [
        ("-c example.config -p ev6_3D.ptrace -grid_layer_file ev6_3D.lcf -model_type grid -detailed_3D on "
         "-grid_steady_file example.grid.steady -steady_file example.steady",
         ["example.steady", "example.grid.steady"]),
        ("-c example.config -p ev6_3D.ptrace -grid_layer_file ev6_3D.lcf -init_file @init -model_type grid "
         "-detailed_3D on -o example.transient", ["example.transient"]),
    ]),
}


# inputs only: examples the reference's default build cannot run (example 5's
# 39 x 39 microchannel stack needs its SuperLU build, absent here); the GPU
# tests compare the command line's outputs on them with the oracle
INPUT_ONLY = {
    "example5_mc": ("example5", ["example.config", "example.lcf", "example.materials", "example.ptrace",
                                 "floorplans/interposer.flp", "floorplans/silicon.flp", "floorplans/heater.flp",
                                 "floorplans/DRAM.flp", "microchannel_geometries/horizontal.csv"]),
}


def copy_inputs(name, spec):
    exdir, inputs = spec
    for f in inputs:
        dst = os.path.join(OUT, name, f)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copy(os.path.join(REF, exdir, f), dst)


# a package file of this build's own: natural convection over a fin-channel sink
NATURAL = "# natural convection, fin-channel sink\n-natural_convec\t1\n-sink_type\t0\n-fin_height\t0.03\n" \
          "-fin_width\t0.001\n-channel_width\t0.002\n"


def input_file(exdir, f, dst):
    """copies an input: 'name' from the example's directory, '@dir/name' from
    another example, '@natural.config' is written here"""
    if f == "@natural.config":
        with open(os.path.join(dst, "natural.config"), "w") as fo:
            fo.write(NATURAL)
    elif f.startswith("@"):
        shutil.copy(os.path.join(REF, f[1:]), os.path.join(dst, os.path.basename(f)))
    else:
        shutil.copy(os.path.join(REF, exdir, f), os.path.join(dst, f))


def run_case(name, spec):
    exdir, inputs, runs = spec
    dst = os.path.join(OUT, name)
    os.makedirs(os.path.join(dst, "expected"), exist_ok=True)
    for f in inputs:
        input_file(exdir, f, dst)
    with tempfile.TemporaryDirectory() as tmp:
        for f in inputs:
            input_file(exdir, f, tmp)
        init = None
        for argv, keep in runs:
            args = argv.split()
            if "@init" in args:
                args[args.index("@init")] = init
            print(name, " ".join(args), flush=True)
            if keep and keep[0].startswith("!"):   # a run that must fail: keep its message
                r = subprocess.run([EXE] + args, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                                   text=True)
                assert r.returncode != 0, "expected the reference to stop with an error"
                with open(os.path.join(dst, "expected", keep[0][1:]), "w") as fo:
                    fo.write(r.stderr.strip().splitlines()[-1] + "\n")
                continue
            subprocess.run([EXE] + args, cwd=tmp, check=True, stdout=subprocess.DEVNULL)
            for k in keep:
                src = os.path.join(tmp, k)
                if ".grid." in k:   # per-cell dumps: kept compressed
                    with open(src, "rb") as fi, gzip.open(os.path.join(dst, "expected", k + ".gz"), "wb") as fo:
                        fo.write(fi.read())
                else:
                    shutil.copy(src, os.path.join(dst, "expected", k))
            if init is None:
                init = keep[0]
                shutil.copy(os.path.join(tmp, init), os.path.join(tmp, init + ".init"))
                init = init + ".init"


def main():
    if not os.path.exists(EXE):
        sys.exit("build the reference first: make -C oracle ref cli")
    only = sys.argv[1:]
    for name, spec in CASES.items():
        if not only or name in only:
            run_case(name, spec)
    for name, spec in INPUT_ONLY.items():
        if not only or name in only:
            copy_inputs(name, spec)


if __name__ == "__main__":
    main()
