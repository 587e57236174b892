"""Build libmcflow.so (HIP, gfx950) in-tree.

    python -m minicombust_app_amd.build            # library
    python -m minicombust_app_amd.build --driver   # also the C++ drop-in driver

The library is compiled with hipcc for gfx950 only; packed FP32 instructions
are disabled (-target-feature -packed-fp32-ops).  Parity-critical source files
turn FP contraction off with a file-scope pragma.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "libmcflow.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(jobs=8):
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(LIB_DIR, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [HIPCC] + COMMON + ["-c", s, "-o", o]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise SystemExit("compile failed")
    for p in procs:
        if p.wait() != 0:
            raise SystemExit("compile failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_driver():
    src = os.path.join(HERE, "host", "minicombust_amd.cpp")
    out = os.path.join(LIB_DIR, "minicombust_amd")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "host"),
          src, "-L" + LIB_DIR, "-Wl,-rpath," + LIB_DIR, "-lmcflow", "-o", out])
    return out


if __name__ == "__main__":
    build_lib()
    if "--driver" in sys.argv:
        build_driver()
