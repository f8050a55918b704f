"""Build libpt_hip.so in place: hipcc for gfx950, one object per source file.

Usage:  python -m particle_transport_amd.build   (or build_library() from Python)

Per-file flags:
  pt_push_ref.hip   -ffp-contract=off  (FP64 parity path: the reference's
                                        arithmetic has no fused multiply-adds)
  pt_inject.hip     -ffp-contract=off  (positions bit-identical to plain FP64)
  pt_stats.hip      -munsafe-fp-atomics (hardware FP64 atomics on coarse-grained
                                        hipMalloc memory)
All files: packed FP32 instructions disabled (-target-feature -packed-fp32-ops).

libpt_calib.so (csrc/pt_calib.hip) holds the vector-ALU issue-rate kernels
bench.py times; it is not part of the boundary.

libpt_h5.so (csrc/pt_h5.cpp, host C++ over the HDF5 C library) is built beside
it with g++: HDF5 from $PT_HDF5_DIR (default /opt/conda), found at run time
through an rpath; libstdc++ is linked statically so that loading it never
pulls a second C++ runtime into the process.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpt_hip.so")
OBJDIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("PT_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function"]
SOURCES = {
    "pt_runtime.hip": [],
    "pt_push_ref.hip": ["-ffp-contract=off"],
    "pt_push_fast.hip": [],
    "pt_stats.hip": ["-munsafe-fp-atomics"],
    "pt_serve.hip": [],
    "pt_sort.hip": [],
    "pt_inject.hip": ["-ffp-contract=off"],
    "pt_cells.hip": [],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: install ROCm or set HIPCC")


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


H5_LIB = os.path.join(HERE, "libpt_h5.so")
HDF5_DIR = os.environ.get("PT_HDF5_DIR", "/opt/conda")


def build_h5(verbose: bool = False) -> str:
    inc = os.path.join(HDF5_DIR, "include")
    lib = os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError(f"hdf5.h not found under {inc}: set PT_HDF5_DIR to an HDF5 1.10+ installation")
    tmp = H5_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-Wextra",
           "-I", os.path.join(os.path.dirname(HERE), "include"), "-I", inc, os.path.join(CSRC, "pt_h5.cpp"),
           "-o", tmp, "-static-libstdc++", "-static-libgcc", "-L", lib, "-lhdf5", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for pt_h5.cpp:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5_LIB)
    if verbose:
        print("built", H5_LIB)
    return H5_LIB


CALIB_LIB = os.path.join(HERE, "libpt_calib.so")


def build_calib(verbose: bool = False) -> str:
    """libpt_calib.so: the vector-ALU issue-rate kernels bench.py times
    (csrc/pt_calib.hip; not part of the libpt_hip.so boundary)."""
    tmp = CALIB_LIB + ".tmp"
    cmd = [hipcc(), *COMMON, "-shared", os.path.join(CSRC, "pt_calib.hip"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for pt_calib.hip:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CALIB_LIB)
    if verbose:
        print("built", CALIB_LIB)
    return CALIB_LIB


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
    build_h5(verbose=True)
    build_calib(verbose=True)
    sys.exit(0)
