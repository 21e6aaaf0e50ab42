"""Builds libstart_gpu.so (gfx950) in-tree with hipcc.

Flags:
  -ffp-contract=off       every + and * rounds separately (FP64 parity with the
                          CPU restatement of the reference arithmetic)
  -munsafe-fp-atomics     hardware FP64 atomic add for the charge scatter
  -packed-fp32-ops off    no packed FP32 instructions are generated
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libstart_gpu.so")
SOURCES = ["start_gpu.hip", "step_kernels.hip", "mesh_prep.hip", "sort.hip", "sources.hip"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


OBJ = os.path.join(HERE, "_build")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def build_library(verbose: bool = False) -> str:
    """Every source compiled to an object in parallel (the step kernels are the
    long pole), then one link."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
        cmd = [hipcc()] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-c", "-o", obj,
                                       os.path.join(CSRC, s)]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((s, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for s, obj, p in jobs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(s)
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    cmd = [hipcc(), "--offload-arch=gfx950", "-fPIC", "-shared", "-Wl,--no-undefined", "-o", LIB + ".tmp"] + \
        [obj for _, obj, _ in jobs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libstart_gpu.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libstart_host.so")
IO_LIB = os.path.join(HERE, "libstart_io.so")
SHIM_CHECK = os.path.join(HERE, "shim_check")
H5_LIB = os.path.join(HERE, "libstart_h5.so")
# the image's HDF5 C library (1.10): headers and the shared library
HDF5_PREFIX = "/opt/conda"


def build_host(verbose: bool = False) -> list[str]:
    """The C++ step-loop entry points (libstart_host.so) over libstart_gpu.so,
    the shim_check driver that runs them like the reference's main loop, and
    the Gmsh 4.1 reader and the output writers (libstart_io.so, host only), and
    the HDF5 deposition file (libstart_h5.so, host only, when the HDF5 C
    library is present)."""
    gxx = shutil.which("g++") or "g++"
    common = ["-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), "-I", HOST]
    cmds = [
        [gxx] + common + ["-fopenmp", "-fPIC", "-shared", "-o", HOST_LIB + ".tmp",
                          os.path.join(HOST, "StartProcessors.cpp"), os.path.join(HOST, "DeviceConfig.cpp"), "-L", HERE,
                          "-lstart_gpu", "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined"],
        [gxx] + common + ["-fopenmp", "-o", SHIM_CHECK + ".tmp", os.path.join(HOST, "shim_check.cpp"),
                          "-L", HERE, "-lstart_host", "-lstart_gpu", "-Wl,-rpath,$ORIGIN"],
        [gxx] + common + ["-fPIC", "-shared", "-o", IO_LIB + ".tmp", os.path.join(HOST, "GmshReader.cpp"),
                          os.path.join(HOST, "GmshWriter.cpp"), os.path.join(HOST, "Outputs.cpp"),
                          "-Wl,--no-undefined"],
    ]
    outs = [HOST_LIB, SHIM_CHECK, IO_LIB]
    h5_inc = os.path.join(HDF5_PREFIX, "include")
    h5_lib = os.path.join(HDF5_PREFIX, "lib")
    if os.path.exists(os.path.join(h5_inc, "hdf5.h")) and os.path.exists(os.path.join(h5_lib, "libhdf5.so")):
        gcc = shutil.which("gcc") or "gcc"
        cmds.append([gcc, "-std=c11", "-O2", "-Wall", "-Wextra", "-D_POSIX_C_SOURCE=200809L", "-I",
                     os.path.join(ROOT, "include"), "-I", h5_inc, "-fPIC", "-shared", "-o", H5_LIB + ".tmp",
                     os.path.join(HOST, "TriangleHdf5.c"), "-L", h5_lib, "-lhdf5", "-lm", "-Wl,-rpath," + h5_lib,
                     "-Wl,--no-undefined"])
        outs.append(H5_LIB)
    elif verbose:
        print(f"no HDF5 under {HDF5_PREFIX}: libstart_h5.so not built", flush=True)
    for cmd, out in zip(cmds, outs):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"failed building {os.path.basename(out)}")
        os.replace(out + ".tmp", out)
    return outs


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_host(verbose=True))
