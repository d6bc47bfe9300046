"""CPU restatement of the device initial conditions (sph_init_run). TEST INFRASTRUCTURE ONLY.

Follows packages/pmacc/include/spmacc/particles/initialization/SC.hpp:50-59 (truncateNearInteger),
:118-146 (computeSCCellCounts: x truncated with a 4-epsilon round-off tolerance, other axes rounded
to nearest, at least 1) and :178-189 (cell-centre positions, x fastest), evaluated in float64, plus
the Philox4x32-10 jitter (Salmon, Moraes, Dror & Shaw, SC'11) and the set-up fields. Positions, ids,
masses and internal energies are bit-exact against the device; fields computed through sin() agree
to the last bits of float32.
"""
from __future__ import annotations

import math
import sys

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, key) -> np.ndarray:
    """ctr: uint32 array (4, n); key: (k0, k1). Returns the (4, n) uint32 outputs."""
    c = [ctr[i].astype(np.uint64) for i in range(4)]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return np.stack(c).astype(np.uint32)


def truncate_near_integer(v: float) -> int:
    t = int(v)
    scale = v if v > 1.0 else 1.0
    tol = 4.0 * sys.float_info.epsilon * scale
    dist = float(t + 1) - v
    return t + 1 if 0.0 <= dist <= tol else t


def sc_counts(lo, hi, spacing) -> tuple:
    n0 = truncate_near_integer((hi[0] - lo[0]) / spacing) or 1
    rest = [max(1, int((hi[a] - lo[a]) / spacing + 0.5)) for a in (1, 2)]
    return (max(1, n0), rest[0], rest[1])


def lattice(blocks, seed: int, jitter: float, h0: float, gamma: float):
    """Particles of sph_init_run for capi.LatticeBlock-like objects; arrays keyed by field name."""
    out = {k: [] for k in ("id", "species", "pos", "vel", "B", "mass", "u", "rho")}
    first = 0
    g32 = float(np.float32(gamma))
    for b in blocks:
        lo, hi = list(b.lo), list(b.hi)
        n = sc_counts(lo, hi, b.spacing)
        cnt = n[0] * n[1] * n[2]
        g = np.arange(first, first + cnt, dtype=np.uint64)
        li = np.arange(cnt, dtype=np.int64)
        idx = [li % n[0], (li // n[0]) % n[1], li // (n[0] * n[1])]
        ctr = np.zeros((4, cnt), dtype=np.uint32)
        ctr[0] = (g & MASK32).astype(np.uint32)
        ctr[1] = (g >> np.uint64(32)).astype(np.uint32)
        r = philox4x32_10(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))
        x, d = [], []
        for a in range(3):
            da = (hi[a] - lo[a]) / float(n[a])
            v = lo[a] + (idx[a].astype(np.float64) + 0.5) * da
            u = (r[a].astype(np.float64) + 0.5) * 2.3283064365386963e-10
            v = v + jitter * da * (2.0 * u - 1.0)
            x.append(v)
            d.append(da)
        rho = np.full(cnt, b.density)
        vel = [np.full(cnt, float(b.velocity[a])) for a in range(3)]
        bf = [np.full(cnt, float(b.bfield[a])) for a in range(3)]
        p = list(b.params)
        twopi = 6.283185307179586
        if b.setup == 1:  # rotor
            rx, ry = x[0] - p[0], x[1] - p[1]
            rr = np.sqrt(rx * rx + ry * ry)
            disc = rr < p[2]
            om = p[3] / p[2]
            rho = np.where(disc, p[4], rho)
            vel[0] = np.where(disc, -om * ry, vel[0])
            vel[1] = np.where(disc, om * rx, vel[1])
        elif b.setup == 2:  # Orszag-Tang
            vel[0] = -p[0] * np.sin(twopi * x[1])
            vel[1] = p[0] * np.sin(twopi * x[0])
            bf[0] = -p[1] * np.sin(twopi * x[1])
            bf[1] = p[1] * np.sin(2.0 * twopi * x[0])
        elif b.setup == 3:  # shear layer
            band = np.abs(x[1] - 0.5) < 0.25
            rho = np.where(band, b.density * p[1], b.density)
            vel[0] = np.where(band, 0.5 * p[0], -0.5 * p[0])
            vel[1] = p[2] * np.sin(2.0 * twopi * x[0] + p[3])
        vol = d[0] * d[1] * d[2]
        out["id"].append(g)
        out["species"].append(np.full(cnt, b.species, dtype=np.uint8))
        out["pos"].append(np.stack(x))
        out["vel"].append(np.stack(vel).astype(np.float32))
        out["B"].append(np.stack(bf).astype(np.float32))
        out["mass"].append((rho * vol).astype(np.float32))
        out["u"].append((b.pressure / ((g32 - 1.0) * rho)).astype(np.float32))
        out["rho"].append(rho.astype(np.float32))
        first += cnt
    res = {k: np.concatenate(v, axis=-1) for k, v in out.items()}
    res["h"] = np.full(first, np.float32(h0), dtype=np.float32)
    return res
