"""Reads HDF5 files through tests/cpp/h5_dump_json.cpp (test infrastructure: no h5py in this image).

read(path) -> {object path: entry}; an entry has "kind" ("group" / "dataset"), "attrs"
({name: {"dtype", "value"}}) and, for datasets, "dtype", "shape" and the contents as "array".
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

import numpy as np

from spearhed_amd import build as B

HERE = os.path.dirname(os.path.abspath(__file__))
_EXE = None


def dumper() -> str:
    global _EXE
    if _EXE is not None:
        return _EXE
    found = B.find_hdf5()
    if found is None:
        raise RuntimeError("no HDF5 installation found")
    inc, libdir = found
    exe = os.path.join(tempfile.mkdtemp(prefix="h5dump_"), "h5_dump_json")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{inc}", os.path.join(HERE, "cpp", "h5_dump_json.cpp"),
                    "-o", exe, f"-L{libdir}", "-lhdf5", f"-Wl,-rpath,{libdir}", "-static-libstdc++",
                    "-static-libgcc"], check=True, capture_output=True, text=True)
    _EXE = exe
    return exe


def read(path: str) -> dict:
    out = tempfile.mkdtemp(prefix="h5data_")
    res = subprocess.run([dumper(), path, out], check=True, capture_output=True, text=True, timeout=300)
    tree = json.loads(res.stdout)
    for e in tree.values():
        if e["kind"] == "dataset":
            code = e["dtype"]
            e["array"] = np.fromfile(os.path.join(out, e["data"]), dtype=np.dtype(code)).reshape(e["shape"])
    return tree
