"""End-to-end physics check of the oracle's step on the reference's validation problem (CPU).

The reference validates its Sod set-up by binning particles in x at t = 0.2 and taking L1 errors
against the exact Riemann solution (share/spearhed/ValidationSetups/SodShockTube/validate_sod.py:
270-290; indicative L1 rho ~0.02, vx ~0.03, P ~0.03, u ~0.05 in README.md:93-100). Here the
same measure is taken on a small 3D tube (constant h, frozen wall slabs, Monaghan viscosity) run
with the oracle's full step (drift, resort, density, forces, kick, CFL dt).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ref as O
import riemann
from spearhed_amd import ic


def test_riemann_star_state_sod():
    # Toro (2009) table 4.2, test 1 (Sod): p* = 0.30313, u* = 0.92745
    p, u = riemann.star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    assert p == pytest.approx(0.30313, abs=1e-5)
    assert u == pytest.approx(0.92745, abs=1e-5)


def test_riemann_sample_limits():
    x = np.array([0.0, 0.45, 0.55, 1.0])
    r, u, p = riemann.sample(x, 0.01)
    assert r[0] == 1.0 and p[0] == 1.0 and r[-1] == 0.125 and p[-1] == 0.1


def test_sod_tube_l1(oracle_lib):
    cfg, ps = ic.sod_tube(seed=1, nx_left=64, cross=4)
    o = O.OracleRun(cfg, ps, "plain")
    o.prime()
    t_end = 0.2
    steps = 0
    while o.st.time < t_end and steps < 2000:
        o.step(1)
        steps += 1
    t = o.st.time
    fluid = o.species == 0
    x = o.pos[0][fluid]
    rho = o.rho[fluid].astype(np.float64)
    vx = o.vel[0][fluid].astype(np.float64)
    u = o.u[fluid].astype(np.float64)
    P = (1.4 - 1) * rho * u
    bins = np.linspace(0.1, 0.9, 33)  # away from the walls, as the tube ends are not part of the test
    centres = 0.5 * (bins[1:] + bins[:-1])
    idx = np.digitize(x, bins) - 1
    ok = (idx >= 0) & (idx < len(centres))
    def binned(f):
        s = np.bincount(idx[ok], weights=f[ok], minlength=len(centres))
        c = np.bincount(idx[ok], minlength=len(centres))
        return s / np.maximum(c, 1)
    er, eu, ep = riemann.sample(centres, t)
    l1 = {
        "rho": float(np.mean(np.abs(binned(rho) - er))),
        "vx": float(np.mean(np.abs(binned(vx) - eu))),
        "P": float(np.mean(np.abs(binned(P) - ep))),
    }
    # coarse tube (64 x 4 x 4 left lattice): bounds a few times the reference's indicative values
    assert l1["rho"] < 0.06, l1
    assert l1["vx"] < 0.09, l1
    assert l1["P"] < 0.06, l1
    assert steps > 50


def test_full_size_sod_fixture_is_consistent():
    """The committed 720-step oracle run (tests/golden/make_sod_golden.py): its stored L1 distances
    follow from its stored profiles and time, and it sits near t = 0.2 with every bin populated."""
    import os

    import sod_profiles as S
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "sod_full_720.npz"))
    prof = {k: z[f"prof_{k}"] for k in S.FIELDS}
    l1 = S.l1_exact(prof, float(z["time"]))
    for k in S.FIELDS:
        assert l1[k] == pytest.approx(float(z[f"l1_{k}"]), rel=1e-12)
    assert 0.2 <= float(z["time"]) < 0.21 and int(z["steps"]) == 720
    assert abs(float(z["time_plain"]) - float(z["time"])) < 5e-3 * float(z["time"])
    assert np.all(z["count"] > 0)
    np.testing.assert_allclose(z["centres"], S.CENTRES)
