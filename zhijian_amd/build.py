"""Builds libzhijian_amd.so in-tree for gfx950 (hipcc for the kernels, g++ for
the host tables) and the C++ drop-in test driver.

    python -m zhijian_amd.build [--verbose]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libzhijian_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions (v_pk_*_f32) stay disabled: some of their forms have given
# silent wrong results on this hardware
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", INC, "-I", CSRC, "-Wall", "-Wno-unused-function"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-I", INC, "-I", CSRC, "-Wall",
             "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include"]

HIP_SRCS = ["zj_kernels.hip", "zj_adjoint.hip", "zj_solver.hip", "zj_tri.hip", "zj_visc.hip", "zj_gpu_api.hip"]
CXX_SRCS = ["zj_host.cpp", "zj_frsolver.cpp", "zj_gmsh.cpp", "zj_h5.cpp", "zj_io.cpp"]
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "frsolver_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "frsolver_driver")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))
    if verbose and (r.stdout or r.stderr):
        sys.stderr.write(r.stdout + r.stderr)


def build(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs, procs = [], []
    for src in HIP_SRCS:
        obj = os.path.join(BUILD, src + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        objs.append(obj)
        procs.append(cmd)
    for src in CXX_SRCS:
        obj = os.path.join(BUILD, src + ".o")
        cmd = ["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        objs.append(obj)
        procs.append(cmd)
    # compile in parallel
    running = []
    for cmd in procs:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for cmd, p in running:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out)
        elif verbose and out:
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    # the C++ drop-in driver (links the library by rpath)
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", INC,
          DRIVER_SRC, "-o", DRIVER, "-L", PKG, "-lzhijian_amd",
          "-Wl,-rpath," + os.path.relpath(PKG, os.path.dirname(DRIVER)).join(["$ORIGIN/", ""])], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="--verbose" in sys.argv)
    print(LIB)
