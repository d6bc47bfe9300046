"""Writes the Orszag-Tang long-run fixtures (run: python tests/golden/make_ot_golden.py [name ...]):

  * ot_full_100.npz -- the mixed-precision oracle, 100 steps (~1 min on 8 cores);
  * ot_fp32_200.npz -- the plain FP32 oracle (FP32 absolute positions, the configuration's stated
    precision), 200 steps, the configuration's stated step count (~4 min).

The Orszag-Tang slab at its stated size (ic.orszag_tang(n=512, nz=4, seed=2): 1,048,576 particles,
SPMHD with Borve correction and hyperbolic div-B cleaning, Monaghan viscosity), primed and stepped
with its own CFL dt. Stored: the step count, the time, the FP64 conserved sums, and the summary
statistics of tests/mhd_stats.py (energies, div-B error, 16 x 16 maps of density and B). The GPU test
repeats each run and compares.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import mhd_stats  # noqa: E402
import oracle_ref as O  # noqa: E402
from spearhed_amd import ic  # noqa: E402

RUNS = {"ot_full_100": ("mixed", 100), "ot_fp32_200": ("plain", 200)}


def make(name):
    prec, steps = RUNS[name]
    cfg, ps = ic.orszag_tang(n=512, nz=4, seed=2)
    o = O.OracleRun(cfg, ps, prec)
    o.prime()
    c = None
    for _ in range(steps // 10):
        c = o.step(10)
    st = mhd_stats.stats(o.to_particles())
    out = {"steps": np.int64(steps), "time": np.float64(o.st.time), "n_particles": np.int64(ps.n),
           "mass": np.float64(c["mass"]), "energy": np.float64(c["energy"]),
           "momentum": np.asarray(c["momentum"], dtype=np.float64)}
    out.update({k: (np.float64(v) if np.isscalar(v) else v) for k, v in st.items()})
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: {ps.n} particles, {steps} steps, t = {o.st.time:.6f}, "
          f"e_kin {st['e_kin']:.6f} e_mag {st['e_mag']:.6f} e_th {st['e_th']:.6f} divb_err {st['divb_err']:.3e}")


if __name__ == "__main__":
    for name in sys.argv[1:] or RUNS:
        make(name)
