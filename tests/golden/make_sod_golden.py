"""Writes tests/golden/sod_full_720.npz (run: python tests/golden/make_sod_golden.py; ~3 min).

The mixed-precision oracle's Sod shock tube at the stated spacings (ic.sod_tube(seed=1): 1/256 and
1/128 lattices, 147,456 fluid particles plus frozen wall slabs, Monaghan viscosity), primed and then
stepped 720 times with its own CFL dt (t ~ 0.2). Stored: the step count, the simulated time, the
bin centres, the binned profiles of rho, v_x, P, u with their counts, and their L1 distances to the
exact Riemann solution. The GPU test repeats the run and compares profile by profile. Also stored:
the simulated time of the plain-FP32 oracle's run of the same 720 steps (`time_plain`), against
which the GPU's plain-FP32 run is checked (the two oracle runs end 1.6e-3 apart in time: their dt
sequences differ through the precision of positions and sums).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_ref as O  # noqa: E402
import sod_profiles as S  # noqa: E402
from spearhed_amd import ic  # noqa: E402

STEPS = 720


def main():
    cfg, ps = ic.sod_tube(seed=1)
    o = O.OracleRun(cfg, ps, "mixed")
    o.prime()
    for _ in range(STEPS // 10):
        o.step(10)
    prof, cnt = S.profiles(o.species, o.pos[0], o.rho, o.vel[0], o.u)
    l1 = S.l1_exact(prof, o.st.time)
    out = {"steps": np.int64(STEPS), "time": np.float64(o.st.time), "centres": S.CENTRES, "count": cnt,
           "n_particles": np.int64(ps.n)}
    for k in S.FIELDS:
        out[f"prof_{k}"] = prof[k]
        out[f"l1_{k}"] = np.float64(l1[k])
    p = O.OracleRun(cfg, ps, "plain")
    p.prime()
    for _ in range(STEPS // 10):
        p.step(10)
    out["time_plain"] = np.float64(p.st.time)
    np.savez_compressed(os.path.join(HERE, "sod_full_720.npz"), **out)
    print(f"sod_full_720: {ps.n} particles, {STEPS} steps, t = {o.st.time:.6f} (plain FP32: {p.st.time:.6f}), "
          f"L1 {l1}")


if __name__ == "__main__":
    main()
