"""Writes tests/golden/ot_full_100.npz (run: python tests/golden/make_ot_golden.py; ~1 min).

The mixed-precision oracle's Orszag-Tang slab at its stated size (ic.orszag_tang(n=512, nz=4,
seed=2): 1,048,576 particles, SPMHD with Borve correction and hyperbolic div-B cleaning, Monaghan
viscosity), primed and stepped 100 times with its own CFL dt. Stored: the step count, the time, the
FP64 conserved sums, and the summary statistics of tests/mhd_stats.py (energies, div-B error, 16 x 16
maps of density and B). The GPU test repeats the run and compares.
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

STEPS = 100


def main():
    cfg, ps = ic.orszag_tang(n=512, nz=4, seed=2)
    o = O.OracleRun(cfg, ps, "mixed")
    o.prime()
    c = None
    for _ in range(STEPS // 10):
        c = o.step(10)
    st = mhd_stats.stats(o.to_particles())
    out = {"steps": np.int64(STEPS), "time": np.float64(o.st.time), "n_particles": np.int64(ps.n),
           "mass": np.float64(c["mass"]), "energy": np.float64(c["energy"]),
           "momentum": np.asarray(c["momentum"], dtype=np.float64)}
    out.update({k: (np.float64(v) if np.isscalar(v) else v) for k, v in st.items()})
    np.savez_compressed(os.path.join(HERE, "ot_full_100.npz"), **out)
    print(f"ot_full_100: {ps.n} particles, {STEPS} steps, t = {o.st.time:.6f}, "
          f"e_kin {st['e_kin']:.6f} e_mag {st['e_mag']:.6f} e_th {st['e_th']:.6f} divb_err {st['divb_err']:.3e}")


if __name__ == "__main__":
    main()
