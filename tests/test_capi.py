"""C ABI checks that need no GPU: the library builds for gfx950, loads, and exports every entry point
declared in include/spearhed_sph.h; the ctypes structures match the header's layout."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from spearhed_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spearhed_sph.h")


def declared_entry_points():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|int64_t|const char\*)\s+(sph_\w+)\s*\(", text, re.M)))


def test_header_declares_the_stage_interface():
    names = declared_entry_points()
    for n in ("sph_ctx_create", "sph_upload", "sph_download", "sph_stage_drift", "sph_stage_resort",
              "sph_stage_neighbour_counts", "sph_stage_density", "sph_stage_forces", "sph_stage_integrate",
              "sph_reduce_min_dt", "sph_reduce_conserved", "sph_step", "sph_graph_capture", "sph_graph_launch",
              "sph_autotune", "sph_sync", "sph_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(capi.LIB_PATH), "build the library first (python -m spearhed_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (sph_\w+)", out))
    missing = [n for n in declared_entry_points() if n not in exported]
    assert not missing, missing


def test_ctypes_binding_covers_header():
    assert set(capi.PROTOTYPES) == set(declared_entry_points())


def test_library_loads_without_gpu_calls():
    lib = capi.load()
    for n in declared_entry_points():
        assert hasattr(lib, n)


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof and field offsets from a C compiler reading the header, against the ctypes mirror."""
    structs = {"sph_run_config": capi.RunConfig, "sph_host_soa": capi.HostSoA, "sph_conserved": capi.Conserved,
               "sph_grid_info": capi.GridInfo, "sph_tuning": capi.Tuning, "sph_profile": capi.Profile}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


@pytest.mark.parametrize("field,value", [
    ("h0", float("inf")), ("h0", float("nan")), ("h0", -0.01), ("gamma", float("nan")), ("gamma", 1.0),
    ("cfl", 0.0), ("cfl", float("nan")), ("cfl_force", -1.0), ("cfl_force", float("inf")),
    ("dt_fixed", float("inf")), ("dt_fixed", float("nan")), ("cell_factor", float("nan")),
    ("box_len", (float("inf"), 1.0, 1.0)), ("box_len", (1.0, 0.0, 1.0)), ("box_lo", (0.0, float("nan"), 0.0)),
    ("box_len", (1e12, 1.0, 1.0)),  # 10^13 supercells on one axis: rejected before any integer conversion
])
def test_invalid_run_configurations_are_rejected_on_the_host(field, value):
    lib = capi.load()
    cfg = capi.make_config(h0=0.01)
    assert lib.sph_run_cells(C.byref(cfg)) > 0
    if isinstance(value, tuple):
        getattr(cfg, field)[:] = value
    else:
        setattr(cfg, field, value)
    assert lib.sph_run_cells(C.byref(cfg)) == -1
    assert lib.sph_bytes_per_run(C.byref(cfg), 1000, capi.PREC_MIXED) == -1


def test_bytes_per_run_is_host_only():
    lib = capi.load()
    cfg = capi.make_config(h0=0.01)
    b = lib.sph_bytes_per_run(C.byref(cfg), 100000, capi.PREC_MIXED)
    assert b > 100000 * 100
    bad = capi.make_config(h0=0.0)
    assert lib.sph_bytes_per_run(C.byref(bad), 10, capi.PREC_MIXED) == -1
    # the admission estimate bounds what the run occupies in a pool sized for it alone (the exact
    # carve of sph_pool_bytes), and stays within 1% of it
    for n in (1000, 4096, 100000, 2097152):
        for mhd in (0, 1):
            for prec in (capi.PREC_MIXED, capi.PREC_FP32):
                cfg = capi.make_config(h0=0.01, mhd=mhd)
                est = lib.sph_bytes_per_run(C.byref(cfg), n, prec)
                exact = lib.sph_pool_bytes(n, lib.sph_run_cells(C.byref(cfg)), 1, prec, mhd)
                assert exact <= est <= 1.01 * exact, (n, mhd, prec, est, exact)
