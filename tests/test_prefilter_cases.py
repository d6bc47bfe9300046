"""The GPU parity cases reach every path of the pair passes' candidate pre-filter (CPU only).

The pair kernels keep, for each supercell of at most 64 particles whose 27-cell stencil holds at
most 768 candidates, only the candidates within the support radius H of the bounding box of the
supercell's targets (with a margin of 2^-13 on H^2). The density pass stores the kept list in the
top of the supercell's mask row when it fits above the masks (masks then use at most 7 of a lane's
8 words); otherwise the force pass recomputes it, and culls again when the masks were not kept.
This restates those rules on the oracle's binning and checks that the `crowded` case of
`test_gpu_parity.py` has supercells on each path, so the GPU tests exercise all of them; and, in
FP32 on the same binning, that the box test drops no candidate the cull accepts for some target
(a box without the H reach, or over a subset of the targets, fails it).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ref
from test_gpu_parity import _crowded_case

PRE_CAP = 768
MASK_WORDS = 8
ROW16 = MASK_WORDS * 64 * 2


def _paths(cfg, ps):
    g = oracle_ref.grid(cfg)
    nx, ny, nz = g.n[0], g.n[1], g.n[2]
    key = oracle_ref.bin_keys(cfg, "mixed", ps.pos)
    perm, cs = oracle_ref.sort_perm(key, ps.id, g.ncell)
    pos = ps.pos[:, perm]
    H2 = float(g.H2)
    out = {"off": 0, "stored": 0, "recomputed": 0, "culled": 0}
    for cell in range(g.ncell):
        cx, cy, cz = cell % nx, (cell // nx) % ny, cell // (nx * ny)
        nt = int(cs[cell + 1] - cs[cell])
        if nt == 0:
            continue
        cand = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    c2 = (((cz + dz) % nz) * ny + (cy + dy) % ny) * nx + (cx + dx) % nx
                    cand.append(np.arange(cs[c2], cs[c2 + 1]))
        cand = np.concatenate(cand)
        if nt > 64 or len(cand) > PRE_CAP:
            out["off"] += 1
            continue
        t0 = pos[:, cs[cell]]
        tp = pos[:, cs[cell]:cs[cell + 1]] - t0[:, None]
        tp -= np.round(tp)
        cp = pos[:, cand] - t0[:, None]
        cp -= np.round(cp)
        d = np.maximum(np.maximum(tp.min(1)[:, None] - cp, 0.0), cp - tp.max(1)[:, None])
        kept = int(((d ** 2).sum(0) < H2 * (1.0 + 1.0 / 8192.0)).sum())
        K = 64 // nt
        per_lane = (kept + K - 1) // K
        if per_lane > 32 * (MASK_WORDS - 1):
            out["culled"] += 1
        elif kept <= ROW16 - 1 - 128 * ((per_lane + 31) // 32):
            out["stored"] += 1
        else:
            out["recomputed"] += 1
    return out


def test_crowded_case_reaches_every_prefilter_path():
    cfg, ps = _crowded_case()
    p = _paths(cfg, ps)
    assert p["stored"] > 300 and p["recomputed"] >= 1 and p["culled"] >= 1, p


def _dropped_accepted(cfg, ps):
    """Pairs the cull accepts whose candidate the box test drops, in FP32 arithmetic (products and
    sums rounded separately here; the kernels fuse them in the same order on both sides)."""
    g = oracle_ref.grid(cfg)
    nx, ny, nz = g.n[0], g.n[1], g.n[2]
    key = oracle_ref.bin_keys(cfg, "mixed", ps.pos)
    perm, cs = oracle_ref.sort_perm(key, ps.id, g.ncell)
    pos = ps.pos[:, perm]
    H2 = np.float32(g.H2)
    H2m = np.float32(H2 * np.float32(1.0 + 1.0 / 8192.0))
    bad = checked = 0
    for cell in range(g.ncell):
        cx, cy, cz = cell % nx, (cell // nx) % ny, cell // (nx * ny)
        if cs[cell + 1] == cs[cell]:
            continue
        cand = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    c2 = (((cz + dz) % nz) * ny + (cy + dy) % ny) * nx + (cx + dx) % nx
                    cand.append(np.arange(cs[c2], cs[c2 + 1]))
        cand = np.concatenate(cand)
        t0 = pos[:, cs[cell]]
        tp = pos[:, cs[cell]:cs[cell + 1]] - t0[:, None]
        tp = (tp - np.round(tp)).astype(np.float32)
        cp = pos[:, cand] - t0[:, None]
        cp = (cp - np.round(cp)).astype(np.float32)
        lo, hi = tp.min(1), tp.max(1)
        d = np.maximum(np.maximum(lo[:, None] - cp, np.float32(0)), cp - hi[:, None])
        d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        kept = d2 < H2m
        r = tp[:, :, None] - cp[:, None, :]
        r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2]
        acc = r2 < H2
        bad += int(acc[:, ~kept].sum())
        checked += int(acc.sum())
    return bad, checked


@pytest.mark.parametrize("case", ["crowded", "turbulence"])
def test_box_test_never_drops_an_accepted_pair(case):
    from spearhed_amd import ic
    cfg, ps = _crowded_case() if case == "crowded" else ic.mhd_turbulence(n=16, seed=17)
    bad, checked = _dropped_accepted(cfg, ps)
    assert checked > 10000 and bad == 0, (bad, checked)
