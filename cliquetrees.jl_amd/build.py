"""In-tree build of the native libraries of the package.

  libmfsym.so  host C++ (g++): symbolic analysis, include/mfsym.h
  libmfhip.so  HIP for gfx950 (hipcc): numeric engine, include/mfhip.h

Run ``python cliquetrees.jl_amd/build.py`` or ``__graft_entry__.build()``.
Every build recompiles; the outputs sit next to this file.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("MF_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["mf_api.hip", "mf_factor.hip", "mf_solve.hip", "mf_refine.hip", "mf_selinv.hip",
               "mf_complete.hip", "mf_plan.cpp", "mf_selinv_plan.cpp"]
SYM_SOURCES = ["symbolic.cpp", "mmd.cpp", "mf_plan.cpp", "mf_selinv_plan.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)


def build_sym():
    srcs = [s for s in SYM_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    defs = ["-DMFS_HAVE_MMD"] if "mmd.cpp" in srcs else []
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", *defs, "-o",
          os.path.join(HERE, "libmfsym.so"), *srcs])


def build_hip(jobs=4):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    procs, objs = [], []
    for s in HIP_SOURCES:
        o = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmd = [hipcc, *flags, *lang, "-c", s, "-o", o]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, cwd=CSRC))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("hipcc failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o",
          os.path.join(HERE, "libmfhip.so"), *objs])


def build_all():
    build_sym()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "sym" in what:
        build_sym()
    if "hip" in what:
        build_hip()
    if "all" in what:
        build_all()
