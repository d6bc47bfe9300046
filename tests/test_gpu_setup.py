"""GPU: initial conditions generated on the device (sph_init_run) against the CPU restatement
(oracle/sph_setup_ref.py): ids, species, positions, masses, internal energies, densities and
smoothing lengths bit-exact; velocity and B (through sin) within 2 float32 ulps. A run filled on the
device steps exactly like the same particles uploaded from the host."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import sph_setup_ref as SR  # noqa: E402

from spearhed_amd import capi, setups  # noqa: E402
from spearhed_amd.particles import ParticleSet  # noqa: E402
from spearhed_amd.stepper import Context  # noqa: E402

pytestmark = pytest.mark.gpu

MAKERS = {"brio_wu": setups.brio_wu, "rotor": setups.rotor, "orszag_tang": setups.orszag_tang,
          "kelvin_helmholtz": setups.kelvin_helmholtz}


def _as_particles(ref) -> ParticleSet:
    n = ref["pos"].shape[1]
    ps = ParticleSet(n)
    ps.id = ref["id"].astype(np.uint64)
    ps.species = ref["species"]
    ps.pos = ref["pos"].copy()
    ps.vel = ref["vel"].copy()
    ps.B = ref["B"].copy()
    ps.mass, ps.u, ps.rho, ps.h = ref["mass"], ref["u"], ref["rho"], ref["h"]
    return ps


@pytest.mark.parametrize("name", sorted(MAKERS))
@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_device_fill_matches_restatement(name, precision):
    cfg, blocks, n = MAKERS[name](12000)
    print(f"setup {name} {precision} n={n}", flush=True)
    ref = SR.lattice(blocks, seed=11, jitter=0.05, h0=cfg.h0, gamma=cfg.gamma)
    with Context([cfg], [n], precision=precision) as ctx:
        ctx.init_run(0, blocks, seed=11, jitter=0.05)
        ctx.sync()
        g = ctx.download(0).sorted_by_id()
    assert np.array_equal(g.id, ref["id"]) and np.array_equal(g.species, ref["species"])
    assert np.all(g.mask == 1)
    pos_ref = ref["pos"] if precision == "mixed" else ref["pos"].astype(np.float32).astype(np.float64)
    assert np.array_equal(g.pos, pos_ref)
    for f in ("mass", "u", "rho", "h"):
        assert np.array_equal(getattr(g, f), ref[f]), f
    for f in ("vel", "B"):
        a, b = getattr(g, f), ref[f]
        assert np.all(np.abs(a - b) <= 2 * np.spacing(np.maximum(np.abs(b), np.float32(1e-30)))), f


def test_device_fill_steps_like_upload():
    cfg, blocks, n = setups.orszag_tang(9000)
    ref = SR.lattice(blocks, seed=5, jitter=0.05, h0=cfg.h0, gamma=cfg.gamma)
    outs = []
    for device in (True, False):
        with Context([cfg], [n], precision="mixed") as ctx:
            if device:
                ctx.init_run(0, blocks, seed=5, jitter=0.05)
                ctx.sync()
                ps = ctx.download(0)  # the device's own fields, so both runs start identical
            else:
                ps = _as_particles(ref)
                ps.vel, ps.B = outs[0][2].vel, outs[0][2].B  # sin() may differ in the last bit
                ctx.upload(0, ps)
            ctx.step(4)
            ctx.sync()
            outs.append((ctx.download(0).sorted_by_id(), ctx.conserved()[0], ps.sorted_by_id()))
    (a, ca, _), (b, cb, _) = outs
    for f in ("pos", "vel", "u", "rho", "B", "psi", "dvdt"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert ca == cb


def test_device_fill_rejects_bad_requests():
    cfg, blocks, n = setups.rotor(4000)
    with Context([cfg], [n + 1], precision="mixed") as ctx:
        with pytest.raises(capi.SphError):
            ctx.init_run(0, blocks)               # size differs from the sites
    with Context([cfg], [n], precision="mixed") as ctx:
        with pytest.raises(capi.SphError):
            ctx.init_run(0, blocks, jitter=0.6)   # jitter must stay inside the cell
        outside = [capi.make_block((0, 0, 0), (1.5, 1.0, blocks[0].hi[2]), blocks[0].spacing, 1.0, 1.0)]
        with pytest.raises(capi.SphError):
            ctx.init_run(0, outside)
