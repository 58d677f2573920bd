"""Build of the gfx950 library letzelphc_amd/liblelph_hip.so (in-tree).

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions
are disabled (-packed-fp32-ops) for every HIP translation unit.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblelph_hip.so")
ARCH = os.environ.get("LELPH_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["lelph_hip.hip", "learned.hip"]
C_SOURCES = ["driver.c", "qeread.c", "asr.c", "phonon.c"]
HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return h


NC4_LIB = os.path.join(HERE, "liblelph_nc4.so")
HDF5_HOME = os.environ.get("LELPH_HDF5_HOME", "/opt/conda")


def build_nc4(verbose: bool = False) -> str | None:
    """The host-only netCDF-4 writer (csrc/nc4w.c) against HDF5 and its
    high-level library; returns None when HDF5 is not installed."""
    inc = os.path.join(HDF5_HOME, "include")
    lib = os.path.join(HDF5_HOME, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5_hl.h")):
        return None
    tmp = NC4_LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-Wall", f"-I{inc}",
           os.path.join(CSRC, "nc4w.c"), "-o", tmp, f"-L{lib}", f"-Wl,-rpath,{lib}",
           "-lhdf5_hl", "-lhdf5"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, NC4_LIB)
    return NC4_LIB


def build(verbose: bool = False) -> str:
    """Compiles every HIP and C source of the library; returns the .so path."""
    build_nc4(verbose)
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = [_hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = ["gcc", "-O2", "-fPIC", "-std=c99", "-Wall", "-c", os.path.join(CSRC, src),
               "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
