"""The C++ host layer (include/spearhed_sph.hpp) compiled with g++ against the built library.

CPU: error reporting and sizing helpers (no device calls). GPU: the stage functors and sph_step give
bit-identical states."""
from __future__ import annotations

import os
import subprocess

import pytest

from spearhed_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = str(tmp_path / "host_stepper")
    lib_dir = os.path.dirname(capi.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "host_stepper.cpp"), "-L", lib_dir, "-lspearhed_sph",
                    f"-Wl,-rpath,{lib_dir}", "-o", exe], check=True)
    return exe


def test_cpp_host_compiles_and_reports_errors(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert '"cpu": "ok"' in out.stdout


@pytest.mark.gpu
def test_cpp_stage_functors_equal_fused_step(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
