"""Builds the product library ``libhotspot_amd.so`` in-tree with hipcc for
gfx950 (one shared object: host C++ model + HIP kernels + the C ABI)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhotspot_amd.so")

SOURCES = ["hs_device.hip", "hs_host.cpp", "hs_config.cpp", "hs_package.cpp", "hs_block.cpp",
           "hs_block.hip"]
DEPS = ["hs_internal.h", "hs_kernels.h", "hs_engine.inl", "hs_batch.inl",
        os.path.join("..", "..", "include", "hotspot_amd.h")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True, jobs: int = 2) -> str:
    """Compiles every source and links the library; returns its path."""
    objdir = os.path.join(ROOT, "_build", "hotspot_amd")
    os.makedirs(objdir, exist_ok=True)
    common = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
              "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]
    # packed FP32 instructions stay off: scalar FP32 only
    gpu = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-munsafe-fp-atomics"]
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        cmd = [hipcc()] + common + (gpu if src.endswith(".hip") else ["-x", "c++"]) + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("compile failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-Wl,-Bsymbolic",
           "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


CLI = os.path.join(HERE, "hotspot")


def build_cli(verbose: bool = True) -> str:
    """Links the command-line simulator (cli/hotspot.cpp) against the library,
    found next to it at run time."""
    tmp = CLI + ".tmp"
    cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", "-x", "c++", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "cli", "hotspot.cpp"), "-L" + HERE, "-lhotspot_amd",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
