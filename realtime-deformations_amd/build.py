"""Build libmpm_hip.so (gfx950) in-tree with hipcc.

    python realtime-deformations_amd/build.py [--jobs N]

Every translation unit is compiled with contraction disabled
(-ffp-contract=off: the deformation-gradient update must round exactly like
the reference's FP32 CPU path; hot loops request FMA explicitly with fmaf) and
without packed FP32 instructions.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "libmpm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "mpm_kernels.hip",
    "mpm_fused.hip",
    "mpm_bin.hip",
    "mpm_capi.hip",
    "mpm_adjoint.hip",
    "mpm_gen.hip",
    "mpm_io.hip",
]
FACADE = os.path.join(PKG, "facade")
DEMO = os.path.join(LIBDIR, "facade_demo")
HOSTCHECK = os.path.join(LIBDIR, "libmpm_hostcheck.so")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    cmds = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            sys.stdout.write(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # the drop-in C++ API compiled with a caller (the reference demo's loop)
    tmp = DEMO + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
          os.path.join(FACADE, "material_point_method.cpp"), os.path.join(FACADE, "demo_main.cpp"),
          "-L" + LIBDIR, "-lmpm_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp])
    os.replace(tmp, DEMO)
    # host build of the device math, for the CPU checks against the oracle
    tmp = HOSTCHECK + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-I" + CSRC, os.path.join(CSRC, "mpm_hostcheck.cpp"), "-o", tmp])
    os.replace(tmp, HOSTCHECK)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose))
