"""Builds libtcamd.so (HIP, gfx950) and the Fortran host tc_host in-tree.

Usage: python taylor-couette-afid_amd/build.py [--jobs N] [--no-fortran]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtcamd.so")
HOST = os.path.join(HERE, "tc_host")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("TC_FC", "/opt/rocm/bin/amdflang")
SOURCES = ["tc_grid.cpp", "tc_abi.cpp", "tc_batch.cpp", "tc_kernels.hip", "tc_poisson.hip",
           "tc_stats.hip", "tc_regrid.hip"]
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed-FP32 instructions anywhere (all arithmetic here is FP64)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-function", "-Wno-unused-value",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_lib(jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    objs, cmds = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmds.append([HIPCC] + HIPFLAGS + lang + ["-c", src, "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                sys.stderr.write(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_host():
    """Fortran host program (bou.in driver with the AFiD routine names)."""
    src = os.path.join(HERE, "fortran", "tc_host.F90")
    if not os.path.exists(src) or not shutil.which(FLANG) and not os.path.exists(FLANG):
        return None
    os.makedirs(BUILD, exist_ok=True)
    tmp = HOST + ".tmp"
    _run([FLANG, "-cpp", "-O2", "-fdefault-real-8", "-fdefault-double-8",
          "-module-dir", BUILD, src, "-o", tmp, "-L" + HERE, "-ltcamd",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST)
    return HOST


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--no-fortran", action="store_true")
    a = ap.parse_args()
    print(build_lib(a.jobs))
    if not a.no_fortran:
        h = build_host()
        if h:
            print(h)


if __name__ == "__main__":
    main()
