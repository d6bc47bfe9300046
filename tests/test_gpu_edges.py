"""GPU edge cases through the C ABI, against the oracle and against the same run alone.

* Every pair inside one interaction radius is found once per direction, dead slots skipped
  (tests/spmacc/unit/ParticleParticleInteraction/ParticleParticleInteraction.cpp:134-208 with
  N = 65 + 130): one supercell holding 195 particles, which takes several target groups per wave.
* Runs with nothing to do (FindNeighbourRegions.cpp:188-213 empty target/source sets,
  SimulationTimestep.cpp:332-398 empty set-ups): a run of frozen walls and tracers has no target,
  so no CFL bound (dt_next = +inf) and does not advance; batched with a fluid run, that run is
  bit-identical to itself alone.
* Tracers under a fixed dt drift by exactly v dt, as the oracle.
* Positions outside a non-periodic box bin to the edge supercells: keys, permutation and cell
  starts bit-exact against the oracle.
* The autotuner's search leaves the context's state as it found it.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_ref
from parity import compare_forces, gpu_stages, oracle_stages
from spearhed_amd import capi, ic
from spearhed_amd.particles import ParticleSet
from spearhed_amd.stepper import Context

pytestmark = pytest.mark.gpu

PRECS = ["plain", "mixed"]
GPREC = {"plain": "fp32", "mixed": "mixed"}


def _one_radius_case(n_dead=15, seed=5):
    rng = np.random.default_rng(seed)
    n = 65 + 130
    ps = ParticleSet(n)
    ps.id[:] = np.arange(n, dtype=np.uint64)
    ps.pos[:] = rng.uniform(-0.2, 0.2, size=(3, n))
    ps.mass[:] = 1.0 / n
    ps.h[:] = 0.5
    ps.rho[:] = 1.0
    ps.u[:] = 1.5
    ps.vel[:] = rng.normal(0.0, 0.1, size=(3, n)).astype(np.float32)
    ps.B[:] = np.stack([np.full(n, 0.1), rng.normal(0, 0.02, n), np.zeros(n)]).astype(np.float32)
    ps.mask[rng.choice(n, n_dead, replace=False)] = 0
    cfg = capi.make_config(box_lo=(-1, -1, -1), box_len=(2, 2, 2), periodic=(0, 0, 0), kernel=capi.KERNEL_CUBIC,
                           h0=0.5, mhd=1, cleaning=1, borve=1)
    return cfg, ps


@pytest.mark.parametrize("prec", PRECS)
def test_all_pairs_in_one_radius_with_dead_slots(prec):
    cfg, ps = _one_radius_case()
    assert oracle_ref.grid(cfg).ncell == 1
    print(f"one radius: n={ps.n} dead={int((ps.mask == 0).sum())} prec={prec}", flush=True)
    g = gpu_stages(cfg, ps, GPREC[prec])
    live = ps.mask[np.argsort(ps.id)] == 1
    n_live = int(live.sum())
    assert np.array_equal(g.nnb, np.where(live, n_live - 1, 0).astype(np.uint32))
    o = oracle_stages(cfg, ps, prec)
    compare_forces(g, o, mhd=True)


def _walls_and_tracers(n_wall=300, n_tracer=40, seed=3):
    rng = np.random.default_rng(seed)
    n = n_wall + n_tracer
    ps = ParticleSet(n)
    ps.id[:] = np.arange(n, dtype=np.uint64)
    ps.pos[:] = rng.uniform(0.02, 0.98, size=(3, n))
    ps.mass[:] = 1e-3
    ps.h[:] = 0.06
    ps.rho[:] = 1.0
    ps.u[:] = 1.5
    ps.vel[:] = rng.normal(0.0, 0.3, size=(3, n)).astype(np.float32)
    ps.species[:n_wall] = capi.SPECIES_BOUNDARY
    ps.species[n_wall:] = capi.SPECIES_TRACER
    return ps


def _cfg_box(dt_fixed=0.0, mhd=1):
    return capi.make_config(box_lo=(0, 0, 0), box_len=(1, 1, 1), periodic=(1, 1, 1), kernel=capi.KERNEL_WENDLAND,
                            h0=0.06, mhd=mhd, cleaning=mhd, dt_fixed=dt_fixed)


@pytest.mark.parametrize("prec", PRECS)
def test_run_without_targets_stays_put_and_leaves_its_batch_alone(prec):
    cfg0, ps0 = _cfg_box(), _walls_and_tracers()
    cfg1, ps1 = ic.mhd_turbulence(n=10, seed=7)
    print(f"no-target run n={ps0.n} batched with turbulence n={ps1.n} prec={prec}", flush=True)
    with Context([cfg0, cfg1], [ps0.n, ps1.n], precision=GPREC[prec]) as ctx:
        ctx.upload(0, ps0)
        ctx.upload(1, ps1)
        ctx.step(4)  # the priming step and three steps
        ctx.sync()
        c = ctx.conserved()
        g0 = ctx.download(0).sorted_by_id()
        g1 = ctx.download(1).sorted_by_id()
    with Context([cfg1], [ps1.n], precision=GPREC[prec]) as ctx:
        ctx.upload(0, ps1)
        ctx.step(4)
        ctx.sync()
        a1 = ctx.download(0).sorted_by_id()
        ca = ctx.conserved()[0]
    want = ps0.pos.astype(np.float32).astype(np.float64) if prec == "plain" else ps0.pos
    assert np.array_equal(g0.pos, want), "a run without a CFL bound must not move"
    assert math.isinf(c[0]["dt_next"]) and c[0]["dt_next"] > 0
    assert c[0]["time"] == 0.0 and c[0]["mass"] == 0.0
    assert int(g0.nnb.sum()) == 0
    for f in ("pos", "vel", "rho", "u", "B", "psi", "nnb", "dvdt"):
        assert np.array_equal(getattr(g1, f), getattr(a1, f)), f
    assert c[1]["time"] == ca["time"] and c[1]["energy"] == ca["energy"]


@pytest.mark.parametrize("prec", PRECS)
def test_tracers_drift_like_the_oracle_with_a_fixed_dt(prec):
    cfg, ps = _cfg_box(dt_fixed=1e-3, mhd=0), _walls_and_tracers(seed=11)
    with Context([cfg], [ps.n], precision=GPREC[prec]) as ctx:
        ctx.upload(0, ps)
        ctx.step(3)  # the priming step and two steps
        ctx.sync()
        g = ctx.download(0).sorted_by_id()
        c = ctx.conserved()[0]
    o = oracle_ref.OracleRun(cfg, ps, prec)
    o.prime()
    o.step(2)
    assert np.array_equal(g.pos, o.pos[:, o.by_id()])
    assert c["time"] == pytest.approx(2e-3, rel=1e-12)


@pytest.mark.parametrize("prec", PRECS)
def test_sort_with_positions_outside_a_nonperiodic_box(prec):
    rng = np.random.default_rng(17)
    n = 2000
    ps = ParticleSet(n)
    ps.id[:] = rng.permutation(n).astype(np.uint64)
    ps.pos[:] = rng.uniform(0.0, 1.0, size=(3, n))
    out = rng.choice(n, 120, replace=False)
    ps.pos[0, out[:60]] = rng.uniform(-0.3, -1e-9, 60)
    ps.pos[0, out[60:]] = rng.uniform(1.0, 1.3, 60)
    ps.pos[2, out[::7]] = 1.0  # exactly on the upper face
    ps.mass[:] = 1.0 / n
    ps.h[:] = 0.05
    ps.rho[:] = 1.0
    ps.u[:] = 1.0
    cfg = capi.make_config(box_lo=(0, 0, 0), box_len=(1, 1, 1), periodic=(0, 1, 0), kernel=capi.KERNEL_WENDLAND,
                           h0=0.05, mhd=0)
    with Context([cfg], [ps.n], precision=GPREC[prec]) as ctx:
        ctx.upload(0, ps)
        ctx.resort()
        key, cs = ctx.sort_info(0)
        g = ctx.download(0)
    okey = oracle_ref.bin_keys(cfg, prec, ps.pos)
    perm, ocs = oracle_ref.sort_perm(okey, ps.id, oracle_ref.grid(cfg).ncell)
    assert np.array_equal(cs, ocs), "cell starts differ"
    assert np.array_equal(g.id, ps.id[perm]), "sort permutation differs"
    assert np.array_equal(key, okey[perm]), "supercell keys differ"


def test_autotune_leaves_the_state_untouched():
    """sph_autotune times candidate launch shapes on the context's own data and restores it: the
    state, the run's clock and dt after the search are those before it, and the winner is applied."""
    cfg, ps = ic.mhd_turbulence(n=12, seed=29)
    with Context([cfg], [ps.n], precision="mixed") as ctx:
        ctx.upload(0, ps)
        ctx.step(3)
        ctx.sync()
        before = ctx.download(0).sorted_by_id()
        c_before = ctx.conserved()[0]
        tuned = ctx.autotune(1)
        after = ctx.download(0).sorted_by_id()
        c_after = ctx.conserved()[0]
        assert ctx.tuning == tuned
    for f in ("id", "pos", "vel", "rho", "u", "B", "psi", "h", "mass"):
        assert np.array_equal(getattr(before, f), getattr(after, f)), f
    for k in ("time", "dt_next", "mass", "energy"):
        assert c_before[k] == c_after[k], k
