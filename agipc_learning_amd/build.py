"""Builds the native libraries in-tree.

  lib/libmasx.so      HIP kernels + host orchestrator + C ABI, gfx950
  lib/libmasx_gen.so  synthetic Newton-system generator (CPU, OpenMP)

Every call compiles (used by __graft_entry__.build()); `ensure()` compiles
only what is missing and is what the Python front end calls on import.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MASX_ARCH", "gfx950")

MASX_SOURCES = ["kernels.hip", "converter.hip", "hierarchy.hip", "fem.hip", "morton.hip", "bandwidth.hip", "devplan.hip",
                "plan.cpp", "solver.cpp", "solver_devplan.cpp", "masx_abi.cpp"]
GEN_SOURCES = ["gen/masx_gen.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_masx(jobs: int = 4):
    os.makedirs(LIB, exist_ok=True)
    objs = []
    procs = []
    common = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
    for src in MASX_SOURCES:
        obj = os.path.join(LIB, "obj_" + os.path.basename(src) + ".o")
        objs.append(obj)
        if src.endswith(".hip"):
            cmd = [HIPCC, f"--offload-arch={ARCH}", *common, "-Xclang", "-target-feature", "-Xclang",
                   "-packed-fp32-ops", "-c", os.path.join(CSRC, src), "-o", obj]
        else:  # host code; the structure analysis uses OpenMP
            cmd = [HIPCC, f"--offload-arch={ARCH}", *common, "-fopenmp", "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("compilation failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compilation failed")
    out = os.path.join(LIB, "libmasx.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return out


def build_gen():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmasx_gen.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o", out + ".tmp",
          *[os.path.join(CSRC, s) for s in GEN_SOURCES]])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_gen()
    build_masx()


def ensure():
    if not os.path.exists(os.path.join(LIB, "libmasx_gen.so")):
        build_gen()
    if not os.path.exists(os.path.join(LIB, "libmasx.so")):
        build_masx()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "masx": build_masx, "gen": build_gen}[what]()
