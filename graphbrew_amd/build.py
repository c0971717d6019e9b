"""Builds libgbk.so (HIP kernels + C ABI, gfx950) and the native bench drivers.

hipcc cross-compiles for gfx950 without a GPU.  Objects go to
graphbrew_amd/_build/, the library to graphbrew_amd/libgbk.so and the
reference-compatible drivers (pr, pr_spmv, bfs, cc, cc_sv, sssp, bc) to
graphbrew_amd/bin/.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgbk.so")
BIN = os.path.join(PKG, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GBK_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled for every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fopenmp", "-Wall", "-I", os.path.join(ROOT, "include"),
              "-I", HOST]

DRIVERS = ["pr", "pr_spmv", "bfs", "cc", "cc_sv", "sssp", "bc"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_lib(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(OBJ, os.path.basename(s)[:-4] + ".o")
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


HOST_LIB = os.path.join(PKG, "libgbkhost.so")


def build_host_lib(jobs: int = 8) -> str:
    """libgbkhost.so: generator, readers, mappings, sources (include/gbk_host.h)."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in sorted(glob.glob(os.path.join(HOST, "*.cc")))
            if os.path.basename(s)[:-3] not in DRIVERS and not os.path.basename(s).startswith("drv_")]
    cxx = shutil.which("g++") or "g++"
    objs, cmds = [], []
    for s in srcs:
        o = os.path.join(OBJ, "host_" + os.path.basename(s)[:-3] + ".o")
        objs.append(o)
        cmds.append([cxx, *HOST_FLAGS, "-fPIC", "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = HOST_LIB + ".tmp"
    _run([cxx, "-shared", "-fopenmp", *objs, "-o", tmp])
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def build_drivers(jobs: int = 8) -> list:
    """C++17 drivers mirroring bench/bin/<kernel> over libgbk.so + libgbkhost.so."""
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cc")))
    mains = {os.path.basename(s)[:-3] for s in srcs} & set(DRIVERS)
    if not mains:
        return []
    os.makedirs(BIN, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    shared = [s for s in sorted(glob.glob(os.path.join(HOST, "drv_*.cc")))]
    out, cmds = [], []
    for name in sorted(mains):
        exe = os.path.join(BIN, name)
        out.append(exe)
        cmds.append([cxx, *HOST_FLAGS, os.path.join(HOST, name + ".cc"), *shared,
                     "-L", PKG, "-lgbk", "-lgbkhost", "-Wl,-rpath,$ORIGIN/..", "-o", exe])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    return out


def build(jobs: int = 8) -> None:
    build_lib(jobs)
    build_host_lib(jobs)
    build_drivers(jobs)


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print(LIB)
