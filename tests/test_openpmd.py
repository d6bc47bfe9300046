"""The openPMD writer (include/spearhed_openpmd.h) on the CPU: file layout against the openPMD base
standard 1.1.0, record names and units against the reference's plugin, values against the input.

The standard (github.com/openPMD/openPMD-standard, STANDARD.md 1.1.0) requires at the file root the
attributes openPMD ("1.1.0"), openPMDextension (uint32), basePath ("/data/%T/"), iterationEncoding
and iterationFormat, and particlesPath when particles are written; each iteration group carries time,
dt and timeUnitSI; each record carries unitDimension (7 powers of L, M, T, I, theta, N, J) and
timeOffset, each record component unitSI; a constant component is a group with the attributes value
and shape; particle species hold position and positionOffset. The reference's writer
(include/spearhed/plugins/openPMD/OpenPMDOutput.hpp:45-166, ParticleTraits.hpp) names the species
"fluid" and the records id, position, mass, velocity, mass_density, specific_internal_energy, with
unitSI 1 throughout. The files are read back by an independent HDF5 walker (tests/cpp/h5_dump_json.cpp).
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import openpmd_io
from spearhed_amd import build as B
from spearhed_amd import capi
from spearhed_amd.particles import ParticleSet

pytestmark = pytest.mark.skipif(B.find_hdf5() is None or not os.path.exists(capi.OPENPMD_LIB_PATH),
                                reason="no HDF5 installation (the openPMD writer is not built)")

SPH_E_INVALID = -1  # include/spearhed_sph.h

# SI unit dimensions written out from the quantities' definitions (L, M, T, I, theta, N, J)
UNITS = {
    "id": [0, 0, 0, 0, 0, 0, 0],
    "position": [1, 0, 0, 0, 0, 0, 0],          # metre
    "positionOffset": [1, 0, 0, 0, 0, 0, 0],
    "mass": [0, 1, 0, 0, 0, 0, 0],              # kilogram
    "velocity": [1, 0, -1, 0, 0, 0, 0],         # metre / second
    "mass_density": [-3, 1, 0, 0, 0, 0, 0],     # kilogram / metre^3
    "specific_internal_energy": [2, 0, -2, 0, 0, 0, 0],  # joule / kilogram
    "magnetic_field": [0, 1, -2, -1, 0, 0, 0],  # tesla = kg s^-2 A^-1
    "psi": [1, 1, -3, -1, 0, 0, 0],             # tesla * metre / second
}


def _soa(ps: ParticleSet):
    return ps.soa()


def _particles(n=300, seed=1):
    rng = np.random.default_rng(seed)
    ps = ParticleSet(n)
    ps.id[:] = rng.permutation(n).astype(np.uint64) + 1000
    ps.pos[:] = rng.uniform(-1, 1, size=(3, n))
    ps.mass[:] = rng.uniform(0.5, 1.5, n).astype(np.float32)
    ps.vel[:] = rng.normal(size=(3, n)).astype(np.float32)
    ps.rho[:] = rng.uniform(0.1, 2, n).astype(np.float32)
    ps.u[:] = rng.uniform(1, 3, n).astype(np.float32)
    ps.h[:] = 0.1
    ps.B[:] = rng.normal(size=(3, n)).astype(np.float32)
    ps.psi[:] = rng.normal(size=n).astype(np.float32)
    ps.species[:] = rng.choice([0, 0, 0, 1, 2], n).astype(np.uint8)
    ps.mask[:] = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    return ps


def _write(ps, tmp_path, mhd=1, it=42, time=0.125, dt=1e-3, pattern="out_%06T.h5"):
    lib = capi.load_openpmd()
    pat = str(tmp_path / pattern)
    soa = _soa(ps)
    rc = lib.sph_openpmd_write_soa(C.byref(soa), mhd, pat.encode(), it, time, dt)
    return rc, pat


def test_file_name_pattern():
    lib = capi.load_openpmd()
    buf = C.create_string_buffer(64)
    assert lib.sph_openpmd_file_name(b"output_%06T.h5", 42, buf, 64) == len("output_000042.h5")
    assert buf.value == b"output_000042.h5"
    assert lib.sph_openpmd_file_name(b"a/b_%T.h5", 1234567, buf, 64) > 0 and buf.value == b"a/b_1234567.h5"
    assert lib.sph_openpmd_file_name(b"no_token.h5", 1, buf, 64) == -1
    assert lib.sph_openpmd_file_name(b"output_%06T.h5", 1, buf, 8) == -1  # too small a buffer


def test_layout_units_and_values(tmp_path):
    ps = _particles()
    rc, pat = _write(ps, tmp_path)
    assert rc == capi.SPH_OK, capi.load_openpmd().sph_openpmd_last_error()
    fname = str(tmp_path / "out_000042.h5")
    assert os.path.exists(fname)
    t = openpmd_io.read(fname)
    root = t["/"]["attrs"]
    assert root["openPMD"] == {"dtype": "str", "value": "1.1.0"}
    assert root["openPMDextension"]["dtype"] == "u4" and root["openPMDextension"]["value"] == 0
    assert root["basePath"]["value"] == "/data/%T/"
    assert root["particlesPath"]["value"] == "particles/"
    assert root["iterationEncoding"]["value"] == "fileBased"
    assert root["iterationFormat"]["value"] == os.path.basename(pat) == "out_%06T.h5"
    assert root["author"]["value"] == "spearhed" and root["software"]["dtype"] == "str"
    assert re.fullmatch(r"\d{4}-\d{2}-\d{2} \d{2}:\d{2}:\d{2} [+-]\d{4}", root["date"]["value"])
    it = t["/data/42"]["attrs"]
    assert it["time"] == {"dtype": "f8", "value": 0.125}
    assert it["dt"]["value"] == 1e-3 and it["timeUnitSI"]["value"] == 1.0
    species = {p.split("/")[4] for p in t if p.startswith("/data/42/particles/") and p.count("/") >= 4}
    names = {0: "fluid", 1: "boundary", 2: "tracer"}
    for code, name in names.items():
        sel = (ps.mask == 1) & (ps.species == code)
        assert (name in species) == bool(sel.any())
        base = f"/data/42/particles/{name}"
        n = int(sel.sum())
        want = {"id": ps.id[sel], "mass": ps.mass[sel], "mass_density": ps.rho[sel],
                "specific_internal_energy": ps.u[sel], "psi": ps.psi[sel]}
        dtypes = {"id": "u8", "mass": "f4", "mass_density": "f4", "specific_internal_energy": "f4", "psi": "f4"}
        for rec, arr in want.items():
            e = t[f"{base}/{rec}"]
            assert e["kind"] == "dataset" and e["dtype"] == dtypes[rec] and e["shape"] == [n], rec
            assert np.array_equal(e["array"], arr), rec
            assert e["attrs"]["unitDimension"]["value"] == UNITS[rec], rec
            assert e["attrs"]["timeOffset"] == {"dtype": "f4", "value": 0.0}, rec
            assert e["attrs"]["unitSI"]["value"] == 1.0, rec
        for rec, src, code_ in (("position", ps.pos, "f8"), ("velocity", ps.vel, "f4"),
                                ("magnetic_field", ps.B, "f4")):
            g = t[f"{base}/{rec}"]
            assert g["kind"] == "group" and g["attrs"]["unitDimension"]["value"] == UNITS[rec], rec
            for a, ax in enumerate("xyz"):
                e = t[f"{base}/{rec}/{ax}"]
                assert e["dtype"] == code_ and np.array_equal(e["array"], src[a][sel]), (rec, ax)
                assert e["attrs"]["unitSI"]["value"] == 1.0
        g = t[f"{base}/positionOffset"]
        assert g["kind"] == "group" and g["attrs"]["unitDimension"]["value"] == UNITS["positionOffset"]
        for ax in "xyz":
            c = t[f"{base}/positionOffset/{ax}"]
            assert c["kind"] == "group" and c["attrs"]["value"]["value"] == 0.0
            assert c["attrs"]["shape"]["value"] == [n] and c["attrs"]["shape"]["dtype"] == "u8"


def test_hydro_output_has_no_mhd_records(tmp_path):
    ps = _particles(seed=2)
    rc, _ = _write(ps, tmp_path, mhd=0, it=3)
    assert rc == capi.SPH_OK
    t = openpmd_io.read(str(tmp_path / "out_000003.h5"))
    assert not any("magnetic_field" in p or p.endswith("/psi") for p in t)
    assert "/data/3/particles/fluid/mass_density" in t


def test_errors_are_reported(tmp_path):
    lib = capi.load_openpmd()
    ps = _particles(seed=3)
    soa = ps.soa()
    soa.density = None
    assert lib.sph_openpmd_write_soa(C.byref(soa), 0, str(tmp_path / "x_%T.h5").encode(), 0, 0.0, 0.0) == \
        SPH_E_INVALID
    assert b"required" in lib.sph_openpmd_last_error()
    assert _write(ps, tmp_path, pattern="no_token.h5")[0] == SPH_E_INVALID
    assert lib.sph_openpmd_write_soa(C.byref(ps.soa()), 0, b"/nonexistent_dir/x_%T.h5", 0, 0.0, 0.0) == \
        SPH_E_INVALID


def test_library_exports_every_declared_symbol():
    text = open(os.path.join(os.path.dirname(B.HERE), "include", "spearhed_openpmd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|int32_t|const char\*)\s+(sph_\w+)\s*\(", text, re.M))
    out = subprocess.run(["nm", "-D", "--defined-only", capi.OPENPMD_LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (sph_\w+)", out))
    assert declared and declared <= exported, declared - exported
