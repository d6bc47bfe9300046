"""Shared helpers for GPU-vs-oracle parity tests (test infrastructure)."""
from __future__ import annotations

import numpy as np

import oracle_ref
from spearhed_amd.particles import ParticleSet
from spearhed_amd.stepper import Context

# Per-particle tolerance: |gpu - oracle| <= RTOL * S, where S is the sum over pairs of |term| of the
# accumulated field, each dot product inside a term (v.r, B.r) taken at its magnitude before
# cancellation, sum_a |v_a r_a| (the quantity an FP32 evaluation error is bounded by). Density terms
# are all positive, so for rho this is a plain relative error.
RTOL = 1e-5
# Global FP64 sums (mass, momentum, energy): relative to the sum of magnitudes.
RTOL_GLOBAL = 1e-6


def scaled_err(g, o, scale):
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    return np.abs(g - o) / (np.asarray(scale, dtype=np.float64) + 1e-300)


def assert_scaled(g, o, scale, what, rtol=RTOL, mask=None):
    e = scaled_err(g, o, scale)
    if mask is not None:
        e = e[..., mask]
    worst = float(np.max(e)) if e.size else 0.0
    assert worst <= rtol, f"{what}: max scaled error {worst:.3e} > {rtol:.1e}"
    return worst


def gpu_stages(cfg, ps: ParticleSet, precision: str, stages=("resort", "density", "forces"),
               tuning: dict = None) -> ParticleSet:
    with Context([cfg], [ps.n], precision=precision) as ctx:
        if tuning:
            ctx.set_tuning(**tuning)
        ctx.upload(0, ps)
        for s in stages:
            getattr(ctx, s)()
        ctx.sync()
        return ctx.download(0).sorted_by_id()


def oracle_stages(cfg, ps: ParticleSet, prec: str, stages=("resort", "density", "forces")):
    o = oracle_ref.OracleRun(cfg, ps, prec, scales=True)
    for s in stages:
        getattr(o, s)()
    return o


def oracle_by_id(o: oracle_ref.OracleRun):
    order = o.by_id()
    return order


def compare_forces(g: ParticleSet, o: oracle_ref.OracleRun, mhd: bool, rtol=RTOL):
    order = o.by_id()
    tgt = (o.species[order] == 0) & (o.mask[order] == 1)
    out = {}
    out["rho"] = assert_scaled(g.rho, o.rho[order], o.sc_rho[order], "density", rtol, tgt)
    assert np.array_equal(g.nnb, o.nnb[order]), "neighbour counts differ"
    out["dvdt"] = assert_scaled(g.dvdt, o.dvdt[:, order], o.sc_a[order], "dvdt", rtol, tgt)
    out["dudt"] = assert_scaled(g.dudt, o.dudt[order], o.sc_du[order], "dudt", rtol, tgt)
    vs = np.abs(g.vsig[tgt].astype(np.float64) - o.vsig[order][tgt]) / np.abs(o.vsig[order][tgt])
    assert float(np.max(vs)) <= rtol, f"vsig: {np.max(vs)}"
    if mhd:
        out["dbdt"] = assert_scaled(g.dbdt, o.dbdt[:, order], o.sc_db[order], "dBdt", rtol, tgt)
        out["divb"] = assert_scaled(g.divb, o.divb[order], o.sc_divb[order], "divB", rtol, tgt)
        out["dpsidt"] = assert_scaled(g.dpsidt, o.dpsidt[order], o.sc_dpsi[order] + 1e-30, "dpsidt", rtol, tgt)
    return out
