"""Build the gfx950 HIP library (libspearhed_sph.so) in-tree with hipcc.

Run as ``python -m spearhed_amd.build``. The built library sits next to this file, so it travels
with the source tree; nothing is installed or cached elsewhere.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libspearhed_sph.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["sph_kernels.hip", "sph_context.cpp", "sph_probe.hip", "sph_setup.hip", "sph_noalloc.hip"]
ARCH = "gfx950"

# -packed-fp32-ops off: plain float code must not be turned into packed FP32 instructions.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build_library(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 and link the shared library. Returns its path."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, src.rsplit(".", 1)[0] + ".o")
        cmd = [hipcc(), *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
        objs.append(obj)
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", f"-Wl,-soname,{LIB_NAME}", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


OPENPMD_LIB_PATH = os.path.join(HERE, "libspearhed_openpmd.so")


def find_hdf5():
    """(include dir, library dir) of an HDF5 installation with its C headers, or None."""
    for prefix in (os.environ.get("HDF5_DIR"), "/opt/conda", "/usr/local", "/usr"):
        if not prefix or not os.path.exists(os.path.join(prefix, "include", "hdf5.h")):
            continue
        for libdir in ("lib", "lib64", "lib/x86_64-linux-gnu"):
            if os.path.exists(os.path.join(prefix, libdir, "libhdf5.so")):
                return os.path.join(prefix, "include"), os.path.join(prefix, libdir)
    return None


def build_openpmd(verbose: bool = False):
    """Compile the openPMD writer (host C++ over the HDF5 C API) against the step library. Returns
    its path, or None when no HDF5 installation is found (the step library does not need it)."""
    found = find_hdf5()
    if found is None:
        print("HDF5 headers/library not found: libspearhed_openpmd.so not built", flush=True)
        return None
    inc, libdir = found
    tmp = OPENPMD_LIB_PATH + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{inc}",
           os.path.join(CSRC, "sph_openpmd.cpp"), "-o", tmp, f"-L{HERE}", "-lspearhed_sph", f"-L{libdir}",
           "-lhdf5", f"-Wl,-rpath,{libdir}", "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libspearhed_openpmd.so",
           # no shared libstdc++ of its own: the runpath to the HDF5 installation would otherwise
           # bring that installation's (possibly older) libstdc++ ahead of the system's
           "-static-libstdc++", "-static-libgcc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"openPMD writer build failed:\n{res.stderr}")
    os.replace(tmp, OPENPMD_LIB_PATH)
    return OPENPMD_LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
    print(build_openpmd(verbose="-v" in sys.argv))
