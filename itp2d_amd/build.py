"""Builds the gfx950 library itp2d_amd/lib/libitp2d_amd.so with hipcc.

Sources: csrc/kernels.hip, csrc/mixed.hip, csrc/chol.hip, csrc/tridiag.hip (kernels), csrc/reg_sweep_kernels.hip (the
register engine's sweep kernels, compiled once per precision and line length), csrc/capi.hip (C ABI of
include/itp2d.h),
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

SWEEP_LENGTHS = (32, 64, 128, 192, 256, 512, 1024)

# (source, extra flags, object name)
SOURCES = [(os.path.join(HERE, "csrc", "reg_sweep_kernels.hip"), [f"-DITP_SWEEP_REAL={r}", f"-DITP_SWEEP_N={n}"],
            f"reg_sweep_{r}_{n}.o") for r in ("double", "float") for n in SWEEP_LENGTHS] + [
    (src, [], os.path.basename(src) + ".o") for src in [
    os.path.join(HERE, "csrc", "kernels.hip"),
    os.path.join(HERE, "csrc", "mixed.hip"),
    os.path.join(HERE, "csrc", "chol.hip"),
    os.path.join(HERE, "csrc", "tridiag.hip"),
    os.path.join(HERE, "csrc", "capi.hip"),
    os.path.join(HERE, "host", "itpsystem.cpp"),
    os.path.join(HERE, "host", "batch.cpp"),
]]

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


def _newest_input(src):
    # the source and every header it can include
    heads = [os.path.join(d, f) for d in (os.path.join(HERE, "csrc"), os.path.join(HERE, "host"),
                                          os.path.join(ROOT, "include")) for f in os.listdir(d)
             if f.endswith((".hpp", ".h"))]
    return max(os.path.getmtime(f) for f in [src] + heads)


def build(verbose=False, incremental=False):
    """Compiles every translation unit and links the library. incremental=True
    keeps the objects (under lib/obj) and recompiles only those older than
    their source or a header."""
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(LIB_DIR, "obj") if incremental else LIB_DIR
    os.makedirs(obj_dir, exist_ok=True)
    objs, failed = [], []
    # the translation units compile in parallel, at most `jobs` processes at a time
    jobs = max(1, min(int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16))
    running = []
    for src, extra, oname in SOURCES:
        obj = os.path.join(obj_dir, oname)
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > _newest_input(src):
            objs.append(obj)
            continue
        cmd = [HIPCC] + FLAGS + extra + ["-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        while len(running) >= jobs:
            s0, p0 = running.pop(0)
            if p0.wait() != 0:
                failed.append(s0)
        running.append((f"{src} {' '.join(extra)}", subprocess.Popen(cmd)))
        objs.append(obj)
    for s0, p0 in running:
        if p0.wait() != 0:
            failed.append(s0)
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if not incremental:
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
