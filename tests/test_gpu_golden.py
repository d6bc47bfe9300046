"""GPU (through the C ABI) against the committed golden fixtures: FP64-oracle outputs on seeded
small runs of every kernel and problem family.

  * neighbour counts: bit-exact against the counts of the same precision;
  * density, dv/dt, du/dt, dB/dt, div B: mixed path within 1e-5 of the FP32 summation scale of the
    FP64 result; plain FP32 within 5e-5 (its inputs carry the FP32 rounding of absolute positions);
  * prime + 3 steps: mass to round-off, energy within 1e-6 relative, time within 1e-5 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import CASE_NAMES, load_case
from parity import RTOL, RTOL_GLOBAL, gpu_stages
from spearhed_amd.stepper import Context

pytestmark = pytest.mark.gpu

PREC_TOL = {"fp32": 5e-5, "mixed": RTOL}
NNB_KEY = {"fp32": "plain_nnb", "mixed": "mixed_nnb"}


def _err(g, ref, scale):
    return np.abs(np.asarray(g, np.float64) - ref) / (scale + 1e-300)


@pytest.mark.parametrize("name", CASE_NAMES)
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_gpu_stages_match_fixture(name, prec):
    cfg, ps, z = load_case(name)
    print(f"golden {name} {prec} n={ps.n}", flush=True)
    g = gpu_stages(cfg, ps, prec)
    tol = PREC_TOL[prec]
    tgt = (z["species"] == 0) & (z["mask"] == 1)
    assert np.array_equal(g.nnb, z[NNB_KEY[prec]])
    assert _err(g.rho, z["f64_rho"], z["f64_sc_rho"])[tgt].max() <= tol
    assert _err(g.dvdt, z["f64_dvdt"], z["f64_sc_a"])[:, tgt].max() <= tol
    assert _err(g.dudt, z["f64_dudt"], z["f64_sc_du"])[tgt].max() <= tol
    if cfg.mhd:
        assert _err(g.dbdt, z["f64_dbdt"], z["f64_sc_db"])[:, tgt].max() <= tol
        assert _err(g.divb, z["f64_divb"], z["f64_sc_divb"])[tgt].max() <= tol


@pytest.mark.parametrize("name", CASE_NAMES)
def test_gpu_steps_conserve_like_fixture(name):
    cfg, ps, z = load_case(name)
    print(f"golden steps {name} n={ps.n}", flush=True)
    with Context([cfg], [ps.n], precision="mixed") as ctx:
        ctx.upload(0, ps)
        ctx.prime()
        cons = [ctx.conserved()[0]]
        for _ in range(3):
            ctx.step(1)
            cons.append(ctx.conserved()[0])
    m = np.array([c["mass"] for c in cons])
    e = np.array([c["energy"] for c in cons])
    t = np.array([c["time"] for c in cons])
    assert np.allclose(m, z["steps_mass"], rtol=1e-12, atol=0)
    assert np.allclose(e, z["steps_energy"], rtol=RTOL_GLOBAL, atol=0)
    assert np.allclose(t, z["steps_time"], rtol=1e-5, atol=0)
