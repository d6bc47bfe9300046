"""Summary statistics of an Orszag-Tang slab for long-run comparisons (test infrastructure).

Energies (kinetic sum m v^2/2, magnetic sum m B^2/(2 rho), thermal sum m u), the div-B error measure
mean(|div B| h / |B|) of SPMHD practice (Tricco & Price 2012), and 16 x 16 maps in (x, y) of the
mean density and mean B_x, B_y.
"""
from __future__ import annotations

import numpy as np

NB = 16


def stats(ps) -> dict:
    m = ps.mass.astype(np.float64)
    v2 = np.sum(ps.vel.astype(np.float64) ** 2, axis=0)
    B = ps.B.astype(np.float64)
    b2 = np.sum(B ** 2, axis=0)
    rho = ps.rho.astype(np.float64)
    out = {"e_kin": float(np.sum(0.5 * m * v2)), "e_mag": float(np.sum(0.5 * m * b2 / rho)),
           "e_th": float(np.sum(m * ps.u.astype(np.float64))),
           "divb_err": float(np.mean(np.abs(ps.divb.astype(np.float64)) * ps.h / np.sqrt(b2 + 1e-30)))}
    ix = np.clip((ps.pos[0] * NB).astype(int), 0, NB - 1)
    iy = np.clip((ps.pos[1] * NB).astype(int), 0, NB - 1)
    k = ix + NB * iy
    cnt = np.maximum(np.bincount(k, minlength=NB * NB), 1)
    out["map_rho"] = np.bincount(k, weights=rho, minlength=NB * NB) / cnt
    out["map_bx"] = np.bincount(k, weights=B[0], minlength=NB * NB) / cnt
    out["map_by"] = np.bincount(k, weights=B[1], minlength=NB * NB) / cnt
    return out
