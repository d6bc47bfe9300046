"""Binned Sod-tube profiles and their L1 distance to the exact Riemann solution (test infrastructure).

The measure of the reference's validate_sod.py (share/spearhed/ValidationSetups/SodShockTube/
validate_sod.py:270-290): fluid particles binned in x, bin means of rho, v_x, P = (gamma - 1) rho u
and u compared with the exact solution at the bin centres, L1 = mean absolute difference.
"""
from __future__ import annotations

import numpy as np

import riemann

GAMMA = 1.4
BINS = np.linspace(0.1, 0.9, 65)  # away from the walls, whose neighbourhood is not part of the test
CENTRES = 0.5 * (BINS[1:] + BINS[:-1])
FIELDS = ("rho", "vx", "P", "u")


def profiles(species, pos_x, rho, vx, u):
    """Bin means of rho, v_x, P and u over fluid particles (species 0), and the bin counts."""
    fluid = species == 0
    x = pos_x[fluid]
    rho = rho[fluid].astype(np.float64)
    vals = {"rho": rho, "vx": vx[fluid].astype(np.float64), "u": u[fluid].astype(np.float64)}
    vals["P"] = (GAMMA - 1.0) * rho * vals["u"]
    idx = np.digitize(x, BINS) - 1
    ok = (idx >= 0) & (idx < len(CENTRES))
    cnt = np.bincount(idx[ok], minlength=len(CENTRES))
    out = {k: np.bincount(idx[ok], weights=v[ok], minlength=len(CENTRES)) / np.maximum(cnt, 1)
           for k, v in vals.items()}
    return out, cnt


def l1_exact(prof, t):
    er, eu, ep = riemann.sample(CENTRES, t)
    exact = {"rho": er, "vx": eu, "P": ep, "u": ep / ((GAMMA - 1.0) * er)}
    return {k: float(np.mean(np.abs(prof[k] - exact[k]))) for k in FIELDS}
