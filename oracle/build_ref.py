#!/usr/bin/env python3
"""Build recipe for oracle/_ref/libhecmw_ref.so: the reference's own hecmw1
iterative solver, compiled unmodified from the sources where they lie under
the reference checkout, plus the test-only driver oracle/ref_driver.f90.

Test infrastructure only: the product never links or loads this library.

What is compiled
  * every Fortran file of hecmw1/src/common and hecmw1/src/solver except the
    directories the serial iterative path never reaches and that need
    libraries absent from this image (clustermkl, solver_direct_parallel: MPI;
    contact/ and main/hecmw_solver.f90: they use clustermkl);
  * the thirteen C files of hecmw1/src/common and precond/ml that the linked
    Fortran objects call (timer, comm/control init, message table, ML wrapper
    compiled without ML).  Nothing is stubbed: every symbol resolves to
    reference code.
  flags: flang -cpp -DHECMW_SERIAL -O3 -fopenmp (the reference's serial+OpenMP
  configuration; OpenMP with >=2 threads selects the RCM+multicolour SSOR
  ordering, hecmw_precond_SSOR_33.F90:106-127).

Usage: python3 oracle/build_ref.py [--ref /root/reference] [-j 8]
Outputs go to oracle/_ref/ only.  Exits 0 without building when the reference
checkout is absent (the prebuilt library travels with the tree).
"""
import argparse
import concurrent.futures as cf
import glob
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")

EXCLUDE = ("/clustermkl/", "_unused_code", "/solver_direct_parallel/", "/contact/", "/solver/main/")
C_FILES = [
    "common/hecmw_time.c", "common/hecmw_etype.c", "common/hecmw_comm.c", "common/hecmw_control.c",
    "common/hecmw_ctrllex.c", "common/hecmw_path.c", "common/hecmw_log.c", "common/hecmw_util.c",
    "common/hecmw_error.c", "common/hecmw_lib_fc.c", "common/hecmw_msg.c", "common/hecmw_msg_table.c",
    "solver/precond/ml/hecmw_ML_wrapper.c",
]
FFLAGS = ["-cpp", "-DHECMW_SERIAL", "-O3", "-fopenmp", "-fPIC"]


def fortran_sources(src):
    files = []
    for pat in ("common/*.f90", "common/*.F90", "solver/*/*.f90", "solver/*/*.F90",
                "solver/precond/*/*.f90", "solver/precond/*/*.F90"):
        files += glob.glob(os.path.join(src, pat))
    return sorted(f for f in files if not any(e in f for e in EXCLUDE))


def module_levels(files):
    """Group sources into levels so that every module a file uses is built in
    an earlier level (the module files must exist before a dependant compiles)."""
    provides, uses = {}, {}
    for f in files:
        text = open(f, errors="replace").read()
        for m in re.findall(r"^\s*module\s+(\w+)\s*$", text, re.M | re.I):
            provides[m.lower()] = f
        uses[f] = {u.lower() for u in re.findall(r"^\s*use\s+(\w+)", text, re.M | re.I)}
    level = {}

    def lvl(f, stack=()):
        if f in level:
            return level[f]
        deps = [provides[u] for u in uses[f] if u in provides and provides[u] != f and provides[u] not in stack]
        level[f] = 1 + max((lvl(d, stack + (f,)) for d in deps), default=-1)
        return level[f]

    for f in files:
        lvl(f)
    out = {}
    for f, l in level.items():
        out.setdefault(l, []).append(f)
    return [sorted(out[k]) for k in sorted(out)]


def run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), p.stdout))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    src = os.path.join(args.ref, "hecmw1", "src")
    if not os.path.isdir(src):
        print("build_ref: reference checkout not found at %s; keeping prebuilt oracle/_ref" % args.ref)
        return 0
    obj = os.path.join(OUT, "obj")
    mod = os.path.join(OUT, "mod")
    os.makedirs(obj, exist_ok=True)
    os.makedirs(mod, exist_ok=True)
    fsrc = fortran_sources(src)
    levels = module_levels(fsrc)
    objs = []
    with cf.ThreadPoolExecutor(args.j) as ex:
        for files in levels:
            jobs = []
            for f in files:
                o = os.path.join(obj, os.path.splitext(os.path.basename(f))[0] + ".o")
                objs.append(o)
                jobs.append(ex.submit(run, ["flang"] + FFLAGS + ["-module-dir", mod, "-I", mod, "-c", f, "-o", o]))
            for j in jobs:
                j.result()
        cjobs = []
        for c in C_FILES:
            o = os.path.join(obj, "c_" + os.path.splitext(os.path.basename(c))[0] + ".o")
            objs.append(o)
            cjobs.append(ex.submit(run, ["gcc", "-O2", "-fPIC", "-DHECMW_SERIAL", "-w",
                                         "-I", os.path.join(src, "common"),
                                         "-I", os.path.join(src, "solver", "precond", "ml"),
                                         "-c", os.path.join(src, c), "-o", o]))
        for j in cjobs:
            j.result()
    drv_o = os.path.join(obj, "ref_driver.o")
    run(["flang"] + FFLAGS + ["-module-dir", mod, "-I", mod, "-c", os.path.join(HERE, "ref_driver.f90"), "-o", drv_o])
    lib_a = os.path.join(OUT, "libhecmw_ref_objs.a")
    if os.path.exists(lib_a):
        os.remove(lib_a)
    run(["ar", "rcs", lib_a] + objs)
    vers = os.path.join(OUT, "exports.map")
    with open(vers, "w") as fh:
        fh.write("{ global: ref_*; local: *; };\n")
    so = os.path.join(OUT, "libhecmw_ref.so")
    run(["flang", "-shared", "-fopenmp", "-o", so + ".tmp", drv_o, lib_a,
         "-Wl,--no-undefined", "-Wl,--version-script=" + vers])
    os.replace(so + ".tmp", so)
    print("build_ref: built %s from %d Fortran + %d C reference files" % (so, len(fsrc), len(C_FILES)))
    build_dropin(mod, obj, lib_a)
    return 0


def build_dropin(mod, obj, lib_a):
    """Test binary: the reference's objects + the hecmw_solve_33 Fortran module
    + libhecmw33.so, driven by tests/fortran/dropin_driver.f90 (see there)."""
    root = os.path.dirname(HERE)
    libdir = os.path.join(root, "frontistr_amd", "lib")
    if not os.path.exists(os.path.join(libdir, "libhecmw33.so")):
        print("build_ref: libhecmw33.so not built, skipping the drop-in test binary")
        return
    shim = os.path.join(root, "frontistr_amd", "fortran", "hecmw_solver_33.f90")
    drv = os.path.join(root, "tests", "fortran", "dropin_driver.f90")
    shim_o = os.path.join(obj, "hecmw_solver_33.o")
    drv_o = os.path.join(obj, "dropin_driver.o")
    run(["flang"] + FFLAGS + ["-module-dir", mod, "-I", mod, "-c", shim, "-o", shim_o])
    run(["flang"] + FFLAGS + ["-module-dir", mod, "-I", mod, "-c", drv, "-o", drv_o])
    exe = os.path.join(OUT, "dropin_test")
    run(["flang", "-fopenmp", "-o", exe + ".tmp", drv_o, shim_o, lib_a, "-L", libdir, "-lhecmw33",
         "-Wl,-rpath,$ORIGIN/../../frontistr_amd/lib", "-Wl,--no-undefined"])
    os.replace(exe + ".tmp", exe)
    print("build_ref: built %s" % exe)


if __name__ == "__main__":
    sys.exit(main())
