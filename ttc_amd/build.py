"""Builds libttc_hip.so in-tree (ttc_amd/lib/) for gfx950.

  python -m ttc_amd.build [--force] [-j N]

Steps: write the kernel-instance files (gen_instances.py), compile every
HIP source with hipcc for gfx950 and the planner with g++, link one shared
library.  Device code is compiled with -ffp-contract=off (no FMA
contraction, as the reference's CPU oracle) and with packed-fp32
instructions disabled.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

from . import gen_instances

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libttc_hip.so")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "--offload-arch=" + ARCH, "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-I" + INCLUDE, "-I" + CSRC]

HEADERS = ["ttc_kernels.hpp", "ttc_rc.hpp", "ttc_launch.hpp", "ttc_params.h", "ttc_plan.h", "ttc_runtime.h",
           "gen/inst_all.h"]


def _sources():
    inst, _ = gen_instances.generate()
    hip = [os.path.join(CSRC, "ttc_runtime.hip"), os.path.join(CSRC, "ttc_batch.hip")] + inst
    cxx = [os.path.join(CSRC, "ttc_plan.cpp")]
    return hip, cxx


def _newer(src, obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in [src] + deps)


def _compile(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return cmd, p.returncode, p.stdout


def build(force=False, jobs=None, verbose=False):
    """Compile and link; returns the library path.  Raises on failure."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hip, cxx = _sources()
    deps = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(INCLUDE, "ttc_hip.h")]
    cmds, objs = [], []
    for src in hip + cxx:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        if force or _newer(src, obj, deps):
            if src.endswith(".cpp"):
                cmds.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
            else:
                cmds.append([HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", src, "-o", obj])
    jobs = jobs or min(16, os.cpu_count() or 4)
    failed = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, rc, out in ex.map(_compile, cmds):
            if verbose or rc != 0:
                sys.stderr.write(" ".join(cmd) + "\n" + out)
            if rc != 0:
                failed.append(cmd[-1])
    if failed:
        raise RuntimeError("compilation failed: %s" % ", ".join(failed))
    if force or cmds or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        link = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
        p = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if p.returncode != 0:
            sys.stderr.write(" ".join(link) + "\n" + p.stdout)
            raise RuntimeError("link failed")
        os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build(force=a.force, jobs=a.j, verbose=a.v))


if __name__ == "__main__":
    main()
