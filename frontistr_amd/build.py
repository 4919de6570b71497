#!/usr/bin/env python3
"""Build libhecmw33.so (HIP for gfx950 + host C++) in-tree.

    python3 frontistr_amd/build.py [-j N]

Output: frontistr_amd/lib/libhecmw33.so.  hipcc cross-compiles for gfx950, so
this runs on a machine without a GPU.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "include")]
# Packed FP32 instructions are kept out of every device code object.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]
SOURCES = [
    ("kernels.hip", True),
    ("solver.hip", True),
    ("ordering.cpp", False),
    ("gen_hex8.cpp", False),
    ("gen_tet10.cpp", False),
    ("batch.hip", True),
    ("learned.hip", True),
    ("coarse.hip", True),
    ("noalloc.hip", True),
]


def compile_one(src, hip, objdir):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON
    if hip:
        cmd += ["-x", "hip"] + HIP_FLAGS
    else:
        cmd += ["-x", "c++"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), p.stdout))
    return obj


def build(jobs=4, verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s[0], s[1], objdir), SOURCES))
    out = os.path.join(LIBDIR, "libhecmw33.so")
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fopenmp", "-o", out + ".tmp"] + objs + [
        "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), p.stdout))
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=4)
    a = ap.parse_args()
    build(a.j, verbose=True)
    sys.exit(0)
