"""GPU: runs stepped on the device written as openPMD iterations (sph_write_openpmd) and read back
with the independent HDF5 walker: every record equals the downloaded state, species are split as
the header states, and the iteration's time is the run's clock. The Sod measure of the reference's
validate_sod.py (position/x, velocity/x, mass_density, specific_internal_energy of species "fluid")
taken from the file equals the measure taken from the download.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import openpmd_io
import sod_profiles as S
from spearhed_amd import build as B
from spearhed_amd import capi, ic
from spearhed_amd.stepper import Context

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(B.find_hdf5() is None or not os.path.exists(capi.OPENPMD_LIB_PATH),
                                 reason="no HDF5 installation (the openPMD writer is not built)")]


def _species(t, it, name):
    base = f"/data/{it}/particles/{name}"
    if base not in t:
        return None
    rec = {k: t[f"{base}/{k}"]["array"] for k in ("id", "mass", "mass_density", "specific_internal_energy")}
    for v in ("position", "velocity", "magnetic_field"):
        if f"{base}/{v}" in t:
            rec[v] = np.stack([t[f"{base}/{v}/{ax}"]["array"] for ax in "xyz"])
    if f"{base}/psi" in t:
        rec["psi"] = t[f"{base}/psi"]["array"]
    return rec


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_sod_written_and_read_back(tmp_path, precision):
    cfg, ps = ic.sod_tube(seed=3, nx_left=64, cross=8)
    print(f"openPMD Sod n={ps.n} precision={precision}", flush=True)
    with Context([cfg], [ps.n], precision=precision) as ctx:
        ctx.upload(0, ps)
        ctx.step(21)  # the priming step and 20 steps
        ctx.sync()
        fname = ctx.write_openpmd(0, str(tmp_path / "sod_%06T.h5"), 20)
        g = ctx.download(0)
        c = ctx.conserved()[0]
    assert os.path.basename(fname) == "sod_000020.h5"
    t = openpmd_io.read(fname)
    assert t["/data/20"]["attrs"]["time"]["value"] == c["time"]
    assert t["/data/20"]["attrs"]["dt"]["value"] == c["dt_next"]
    for code, name in ((0, "fluid"), (1, "boundary")):
        sel = (g.mask == 1) & (g.species == code)
        r = _species(t, 20, name)
        assert r is not None and np.array_equal(r["id"], g.id[sel]), name
        assert np.array_equal(r["position"], g.pos[:, sel]) and np.array_equal(r["velocity"], g.vel[:, sel])
        assert np.array_equal(r["mass"], g.mass[sel]) and np.array_equal(r["mass_density"], g.rho[sel])
        assert np.array_equal(r["specific_internal_energy"], g.u[sel])
        assert "magnetic_field" not in r
    assert _species(t, 20, "tracer") is None
    # the reference validator's inputs, read from the file, give the same measure as the download
    f = _species(t, 20, "fluid")
    n = f["id"].size
    prof_file, _ = S.profiles(np.zeros(n, np.uint8), f["position"][0], f["mass_density"], f["velocity"][0],
                              f["specific_internal_energy"])
    prof_dl, _ = S.profiles(g.species[g.mask == 1], g.pos[0][g.mask == 1], g.rho[g.mask == 1],
                            g.vel[0][g.mask == 1], g.u[g.mask == 1])
    assert S.l1_exact(prof_file, c["time"]) == S.l1_exact(prof_dl, c["time"])


def test_mhd_run_written_with_its_magnetic_records(tmp_path):
    cfg, ps = ic.mhd_turbulence(n=12, seed=5)
    with Context([cfg], [ps.n], precision="mixed") as ctx:
        ctx.upload(0, ps)
        ctx.step(3)
        ctx.sync()
        fname = ctx.write_openpmd(0, str(tmp_path / "turb_%T.h5"), 2)
        g = ctx.download(0)
    t = openpmd_io.read(fname)
    r = _species(t, 2, "fluid")
    assert np.array_equal(r["id"], g.id)
    assert np.array_equal(r["magnetic_field"], g.B) and np.array_equal(r["psi"], g.psi)
    assert t[f"/data/2/particles/fluid/magnetic_field"]["attrs"]["unitDimension"]["value"] == [0, 1, -2, -1, 0, 0, 0]


def test_simulation_writes_openpmd_every_period(tmp_path):
    """stepper.Simulation with openpmd_period = 2 (the reference's openPMD.period option): after the
    steps whose number is a multiple of 2, one file per run, holding that step's state."""
    from spearhed_amd.stepper import Simulation
    runs = [ic.kelvin_helmholtz(3000, 81), ic.kelvin_helmholtz(4000, 82)]
    with Context([c for c, _ in runs], [p.n for _, p in runs], precision="mixed") as ctx:
        for r, (_, ps) in enumerate(runs):
            ctx.upload(r, ps)
        ctx.prime()
        sim = Simulation(ctx, openpmd_period=2, openpmd_file=str(tmp_path / "kh_%04T.h5"))
        for step in range(1, 5):
            sim.run_one_step(step=step)
        final = [ctx.download(r) for r in range(len(runs))]
    names = sorted(os.path.basename(f) for f in sim.written)
    assert names == ["kh_0002_run0.h5", "kh_0002_run1.h5", "kh_0004_run0.h5", "kh_0004_run1.h5"]
    for r, g in enumerate(final):
        t = openpmd_io.read(str(tmp_path / f"kh_0004_run{r}.h5"))
        f = _species(t, 4, "fluid")
        assert np.array_equal(f["id"], g.id[g.mask == 1]) and np.array_equal(f["position"], g.pos[:, g.mask == 1])
