"""Builds the gfx950 library itp2d_amd/lib/libitp2d_amd.so with hipcc.

Sources: csrc/kernels.hip, csrc/mixed.hip (kernels), csrc/capi.hip (C ABI of include/itp2d.h),
host/itpsystem.cpp (ITP driver, C ABI of include/itp2d_system.h), host/batch.cpp (batches of
instances on worker threads, C ABI of include/itp2d_batch.h).

Also builds the host-only itp2d_amd/lib/libitp2d_datafile.so (host/datafile.cpp, the HDF5
result file of include/itp2d_datafile.h) with g++ against the HDF5 C library under
HDF5_DIR (default /opt/conda), where its headers are present.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libitp2d_amd.so")
DATAFILE_LIB = os.path.join(LIB_DIR, "libitp2d_datafile.so")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")

SOURCES = [
    os.path.join(HERE, "csrc", "kernels.hip"),
    os.path.join(HERE, "csrc", "mixed.hip"),
    os.path.join(HERE, "csrc", "chol.hip"),
    os.path.join(HERE, "csrc", "capi.hip"),
    os.path.join(HERE, "host", "itpsystem.cpp"),
    os.path.join(HERE, "host", "batch.cpp"),
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-pthread",
    # packed FP32 instructions stay disabled in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    objs, procs = [], []
    # the translation units compile in parallel (one process each)
    for src in SOURCES:
        obj = os.path.join(LIB_DIR, os.path.basename(src) + ".o")
        cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_datafile(verbose)
    return LIB


def build_datafile(verbose=False):
    """The HDF5 result-file library; returns its path, or None without HDF5 headers."""
    inc, lib = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"itp2d_amd: no hdf5.h under {inc}; the datafile library is not built", flush=True)
        return None
    tmp = DATAFILE_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", inc, os.path.join(HERE, "host", "datafile.cpp"), "-o", tmp, "-L", lib, "-lhdf5",
           "-Wl,-rpath," + lib]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DATAFILE_LIB)
    return DATAFILE_LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
