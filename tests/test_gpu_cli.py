"""The command-line simulator (hotspot_amd/hotspot) on the reference's
examples 1-4, against the reference's own command-line outputs on the same
inputs (tests/golden/cli, written by tests/golden/make_cli_goldens.py).

Both print temperatures with "%.2f".  Steady: the reference stops its
Gauss-Seidel sweeps at a 1e-6 K change per sweep, a few 1e-4 to 1e-3 K from
the converged state that this library reaches, so a printed value may differ
by one unit of the last digit (0.01 K) and by no more.  Transient: both
start from the reference's printed steady state (its -init_file) and
integrate with the same adaptive RK4, within 1e-8 K of each other, so the
printed traces agree except where a value sits on a rounding boundary:
at most 0.01 K, and at least 99 % of the entries exactly.
"""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI_DIR = os.path.join(ROOT, "tests", "golden", "cli")
EXE = os.path.join(ROOT, "hotspot_amd", "hotspot")
PRINT_UNIT = 0.01

pytestmark = pytest.mark.gpu

# case -> (steady argv, its steady file, transient argv with @init, outputs to compare)
CASES = {
    "ev6": ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type grid "
            "-steady_file gcc.steady -grid_steady_file gcc.grid.steady", ["gcc.steady", "gcc.grid.steady"],
            "-c example.config -init_file @init -f ev6.flp -p gcc.ptrace -materials_file example.materials "
            "-model_type grid -o gcc.ttrace", ["gcc.ttrace"]),
    "lcf_detailed": ("-c example.config -p example.ptrace -grid_layer_file example.lcf -materials_file "
                     "example.materials -model_type grid -detailed_3D on -steady_file example.steady "
                     "-grid_steady_file example.grid.steady", ["example.steady", "example.grid.steady"],
                     "-c example.config -p example.ptrace -grid_layer_file example.lcf -materials_file "
                     "example.materials -model_type grid -detailed_3D on -init_file @init -o example.ttrace "
                     "-grid_transient_file example.grid.ttrace", ["example.ttrace", "example.grid.ttrace"]),
    "ev6_3d": ("-c example.config -p ev6_3D.ptrace -grid_layer_file ev6_3D.lcf -model_type grid -detailed_3D on "
               "-grid_steady_file example.grid.steady -steady_file example.steady",
               ["example.steady", "example.grid.steady"],
               "-c example.config -p ev6_3D.ptrace -grid_layer_file ev6_3D.lcf -init_file @init -model_type grid "
               "-detailed_3D on -o example.transient", ["example.transient"]),
}


def read_text(path):
    if os.path.exists(path):
        return open(path).read()
    with gzip.open(path + ".gz", "rt") as f:
        return f.read()


def numbers(text):
    """Every numeric field of a HotSpot output file, in order (names,
    'Layer n:' and 't = ...' headers give their numbers too: they must
    match exactly)."""
    out = []
    for tok in text.split():
        try:
            out.append(float(tok))
        except ValueError:
            out.append(tok)
    return out


def compare(got_text, want_text, exact_frac):
    got, want = numbers(got_text), numbers(want_text)
    assert len(got) == len(want)
    strs = [(g, w) for g, w in zip(got, want) if isinstance(w, str)]
    assert all(g == w for g, w in strs)
    g = np.array([x for x in got if not isinstance(x, str)])
    w = np.array([x for x in want if not isinstance(x, str)])
    d = np.abs(g - w)
    assert d.max() <= PRINT_UNIT + 1e-9, d.max()
    assert (d == 0).mean() >= exact_frac, (d == 0).mean()
    return d.max(), (d == 0).mean()


def run_cli(argv, cwd):
    print("cli:", argv, flush=True)
    r = subprocess.run([EXE] + argv.split(), cwd=cwd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.parametrize("case", list(CASES))
def test_cli_matches_reference_outputs(gpu, case, tmp_path):
    src = os.path.join(CLI_DIR, case)
    for f in os.listdir(src):
        if os.path.isfile(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), tmp_path / f)
    steady_argv, steady_out, trans_argv, trans_out = CASES[case]
    run_cli(steady_argv, tmp_path)
    for f in steady_out:
        m, ex = compare(open(tmp_path / f).read(), read_text(os.path.join(src, "expected", f)), 0.5)
        print(f, "max |diff|", m, "exact", ex)
    # the transient starts from the reference's printed steady state
    init = "ref.init"
    shutil.copy(os.path.join(src, "expected", steady_out[0]), tmp_path / init)
    run_cli(trans_argv.replace("@init", init), tmp_path)
    for f in trans_out:
        m, ex = compare(open(tmp_path / f).read(), read_text(os.path.join(src, "expected", f)), 0.99)
        print(f, "max |diff|", m, "exact", ex)


# The package model (-package_model_used 1) on example 2: forced convection
# from the reference's example-1 package file (r_convec from the fan and the
# fin-channel sink), and natural convection (r_convec iterated with the
# steady state until it moves by less than 0.01 K/W).  Every printed value
# stays within one digit of the reference's; under natural convection the
# iteration settles on a slightly different r_convec (the reference's steady
# sweeps stop ~1e-3 K early), so fewer prints agree exactly (over 40 %).
PACKAGE_RUNS = [
    ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type grid "
     "-package_model_used 1 -package_config_file package.config -steady_file forced.steady", ["forced.steady"]),
    ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type grid "
     "-package_model_used 1 -package_config_file natural.config -steady_file natural.steady "
     "-grid_steady_file natural.grid.steady", ["natural.steady", "natural.grid.steady"]),
]


def test_cli_package_model_matches_reference(gpu, tmp_path):
    src = os.path.join(CLI_DIR, "ev6_package")
    for f in os.listdir(src):
        if os.path.isfile(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), tmp_path / f)
    for argv, outs in PACKAGE_RUNS:
        run_cli(argv, tmp_path)
        for f in outs:
            frac = 0.4 if f.startswith("natural") else 0.5
            m, ex = compare(open(tmp_path / f).read(), read_text(os.path.join(src, "expected", f)), frac)
            print(f, "max |diff|", m, "exact", ex)


# The block model on the reference's example 1 (its run.sh command lines),
# with the package model (forced convection; natural convection in the
# steady state) and the leakage fixed point.  The steady state comes from
# the reference's LU factors (within 1e-12 K) and the transient is
# bit-identical to the reference's, so every printed value agrees.
# A natural-convection transient that starts at ambient stops in the first
# interval (no convection at a sink at ambient: the LU finds no pivot), in
# both, with the same message.
BLOCK_RUNS = [
    ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
     "-steady_file gcc.steady -o first.ttrace", ["gcc.steady", "first.ttrace"]),
    ("-c example.config -init_file @init -f ev6.flp -p gcc.ptrace -materials_file example.materials "
     "-model_type block -o gcc.ttrace", ["gcc.ttrace"]),
    ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
     "-package_model_used 1 -package_config_file natural.config -steady_file natural.steady",
     ["natural.steady"]),
    ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
     "-package_model_used 1 -package_config_file package.config -steady_file forced.steady",
     ["forced.steady"]),
    ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
     "-leakage_used 1 -steady_file leak.steady", ["leak.steady"]),
]


def test_cli_block_model_matches_reference(gpu, tmp_path):
    src = os.path.join(CLI_DIR, "ev6_block")
    for f in os.listdir(src):
        if os.path.isfile(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), tmp_path / f)
    shutil.copy(os.path.join(src, "expected", "gcc.steady"), tmp_path / "ref.init")
    for argv, outs in BLOCK_RUNS:
        run_cli(argv.replace("@init", "ref.init"), tmp_path)
        for f in outs:
            m, ex = compare(open(tmp_path / f).read(), read_text(os.path.join(src, "expected", f)), 1.0)
            print(f, "max |diff|", m, "exact", ex)
    argv = ("-c example.config -f ev6.flp -p gcc.ptrace -materials_file example.materials -model_type block "
            "-package_model_used 1 -package_config_file natural.config -o natural.ttrace")
    print("cli:", argv, flush=True)
    r = subprocess.run([EXE] + argv.split(), cwd=tmp_path, capture_output=True, text=True, timeout=600)
    want = open(os.path.join(src, "expected", "natural.err")).read().strip()
    assert r.returncode != 0 and want in r.stderr, r.stderr[-2000:]
