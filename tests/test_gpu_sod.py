"""GPU physics validation on the reference's own validation problem: the Sod shock tube.

The reference validates its Sod set-up by binning particles in x at t = 0.2 and taking L1 errors
against the exact Riemann solution (share/spearhed/ValidationSetups/SodShockTube/validate_sod.py:
270-290), with indicative L1 rho ~0.02, vx ~0.03, P ~0.03, u ~0.05 (README.md:93-100). Here the
whole GPU step (drift, re-sort, density, forces, kick, CFL dt) runs the tube at the stated spacings
(1/256 and 1/128, 147,456 fluid particles, frozen wall slabs, Monaghan viscosity) for 720 CFL steps
(t ~ 0.2):
  * mixed precision against the oracle's run of the same 720 steps (tests/golden/sod_full_720.npz):
    simulated time and every binned profile; plain FP32 against the plain-FP32 oracle's simulated
    time (the two precisions' oracle runs themselves end 1.6e-3 apart in time);
  * both precision modes against the exact solution, within 2.5x the reference's indicative L1
    (its set-up is inviscid; this one carries Monaghan viscosity, which smears the contact and the
    shock over a few h: the oracle's own run gives rho 0.039, vx 0.059, P 0.030, u 0.083).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import sod_profiles as S
from spearhed_amd import ic
from spearhed_amd.stepper import Context

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sod_full_720.npz")
L1_MAX = {"rho": 0.05, "vx": 0.075, "P": 0.075, "u": 0.125}
STEPS = 720
# Largest bin-mean difference allowed against the oracle's run after 720 steps (profiles of order
# one; the runs' dt sequences drift apart by ~1e-4 relative through FP32 summation order, and a few
# dozen particles change bins). Measured: rho 1.0e-4, vx 1.2e-3, P 1.6e-4, u 5.8e-4.
BIN_TOL = {"rho": 5e-4, "vx": 5e-3, "P": 8e-4, "u": 3e-3}


def _run(precision):
    cfg, ps = ic.sod_tube(seed=1)
    print(f"Sod n={ps.n} precision={precision}", flush=True)
    with Context([cfg], [ps.n], precision=precision) as ctx:
        ctx.upload(0, ps)
        ctx.prime()
        for _ in range(STEPS // 10):
            ctx.step(10)
        ctx.sync()
        t = ctx.conserved()[0]["time"]
        g = ctx.download(0)
    prof, cnt = S.profiles(g.species, g.pos[0], g.rho, g.vel[0], g.u)
    return prof, cnt, t


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_sod_full_size_against_the_oracle_and_the_exact_solution(precision):
    prof, cnt, t = _run(precision)
    l1 = S.l1_exact(prof, t)
    print(f"Sod {precision}: t = {t:.6f}, L1 {l1}", flush=True)
    for k, lim in L1_MAX.items():
        assert l1[k] < lim, (k, l1)
    z = np.load(GOLDEN)
    t_ref = float(z["time"] if precision == "mixed" else z["time_plain"])  # the oracle in the same precision
    print(f"Sod {precision}: time {t:.9f} vs the oracle's {t_ref:.9f}", flush=True)
    assert abs(t - t_ref) <= 1e-3 * t_ref
    if precision == "mixed":
        diffs = {k: float(np.max(np.abs(prof[k] - z[f"prof_{k}"]))) for k in S.FIELDS}
        print(f"Sod mixed vs oracle: time {t:.9f} vs {float(z['time']):.9f}, max bin differences {diffs}, "
              f"count differences {int(np.abs(cnt - z['count']).sum())}", flush=True)
        for k, tol in BIN_TOL.items():
            assert diffs[k] <= tol, (k, diffs)
        assert int(np.abs(cnt - z["count"]).sum()) <= 0.002 * int(z["count"].sum())
