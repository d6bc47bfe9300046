"""GPU parity: the HIP path through the C ABI against the CPU oracle on identical seeded inputs.

Contracts (stated tolerances):
  * supercell keys, cell starts, the (cell, id) sort permutation and per-particle neighbour counts:
    bit-exact;
  * density: relative error <= 1e-5;
  * dv/dt, du/dt, dB/dt, dpsi/dt, div B: |gpu - oracle| <= 1e-5 * sum_j |term_ij| (the bound of an
    FP32 summation with cancellation);
  * FP64 global sums (mass, momentum, energy): <= 1e-6 relative to the sum of magnitudes.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ref
from parity import RTOL, RTOL_GLOBAL, compare_forces, gpu_stages, oracle_stages
from spearhed_amd import capi, ic
from spearhed_amd.particles import ParticleSet
from spearhed_amd.stepper import Context

pytestmark = pytest.mark.gpu

PRECS = ["plain", "mixed"]
GPREC = {"plain": "fp32", "mixed": "mixed"}


def _turb(n=16, seed=11):
    return ic.mhd_turbulence(n=n, seed=seed)


@pytest.mark.parametrize("prec", PRECS)
def test_sort_bit_exact(prec):
    cfg, ps = _turb(16, 5)
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


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("problem", ["turb", "kh", "ot_thin"])
def test_forces_parity_mhd(prec, problem):
    if problem == "turb":
        cfg, ps = _turb(16, 3)
    elif problem == "kh":
        cfg, ps = ic.kelvin_helmholtz(6000, 9)
    else:
        cfg, ps = ic.orszag_tang(n=48, nz=4, seed=2)  # one supercell across z: images at +-L
    g = gpu_stages(cfg, ps, GPREC[prec])
    o = oracle_stages(cfg, ps, prec)
    compare_forces(g, o, mhd=True)


@pytest.mark.parametrize("prec", PRECS)
def test_forces_parity_sod_walls(prec):
    cfg, ps = ic.sod_tube(nx_left=32, cross=8, seed=4)
    g = gpu_stages(cfg, ps, GPREC[prec])
    o = oracle_stages(cfg, ps, prec)
    compare_forces(g, o, mhd=False)


@pytest.mark.parametrize("kernel", [capi.KERNEL_CUBIC, capi.KERNEL_QUINTIC, capi.KERNEL_WENDLAND])
def test_forces_parity_kernels(kernel):
    cfg, ps = ic.mhd_turbulence(n=12, seed=21, kernel=kernel)
    g = gpu_stages(cfg, ps, "mixed")
    o = oracle_stages(cfg, ps, "mixed")
    compare_forces(g, o, mhd=True)


def test_density_colocated_kat():
    """DensitySummation.cpp:232-282: N co-located particles have rho = N m W(0, h)."""
    n, h, m = 4, 0.5, 1.0
    cfg = capi.make_config(box_lo=(-1, -1, -1), box_len=(2, 2, 2), periodic=(0, 0, 0), kernel=capi.KERNEL_CUBIC,
                           h0=h, alpha_visc=0.0)
    ps = ParticleSet(n)
    ps.mass[:] = m
    ps.h[:] = h
    ps.rho[:] = 1
    ps.u[:] = 1
    g = gpu_stages(cfg, ps, "fp32", ("resort", "density"))
    expected = n * m * (1.0 / np.pi) / h ** 3
    np.testing.assert_allclose(g.rho, expected, rtol=1e-5)


def test_density_three_in_line_kat():
    """DensitySummation.cpp:285-344: particles at x = 0, 1.5, 3 (h = 1, m = 1.5, cubic)."""
    h, m, dx = 1.0, 1.5, 1.5
    cfg = capi.make_config(box_lo=(-5, -5, -5), box_len=(12, 10, 10), periodic=(0, 0, 0), kernel=capi.KERNEL_CUBIC,
                           h0=h, alpha_visc=0.0)
    ps = ParticleSet(3)
    ps.pos[0] = [0.0, dx, 2 * dx]
    ps.mass[:] = m
    ps.h[:] = h
    ps.rho[:] = 1
    ps.u[:] = 1
    g = gpu_stages(cfg, ps, "fp32", ("resort", "density"))
    w0 = 1.0 / np.pi
    q = dx / h
    w1 = (1.0 / np.pi) * 0.25 * (2 - q) ** 3
    np.testing.assert_allclose(g.rho, [m * (w0 + w1), m * (w0 + 2 * w1), m * (w0 + w1)], rtol=1e-5)


def test_hydro_two_particle_kat():
    """HydroForces.cpp:159-239: two particles at +-0.2 at rest (h 0.5, u = m = 1, gamma 5/3):
    dvdt_x = +-2 m P/rho^2 dW/dr(0.4, 0.5), dvdt_y = dvdt_z = 0, dudt = 0. The GPU stage order always
    runs the density pass first, so rho here is the summed density (identical for both particles)."""
    h, d, gam = 0.5, 0.4, 5.0 / 3.0
    cfg = capi.make_config(box_lo=(-2, -2, -2), box_len=(4, 4, 4), periodic=(0, 0, 0), kernel=capi.KERNEL_CUBIC,
                           h0=h, gamma=gam, alpha_visc=0.0)
    ps = ParticleSet(2)
    ps.pos[0] = [-d / 2, d / 2]
    ps.mass[:] = 1
    ps.h[:] = h
    ps.rho[:] = 1
    ps.u[:] = 1
    g = gpu_stages(cfg, ps, "fp32")
    q = d / h
    w = (1.0 / np.pi) / h ** 3 * (1 - 1.5 * q * q + 0.75 * q ** 3)
    rho = 1.0 * ((1.0 / np.pi) / h ** 3 + w)
    np.testing.assert_allclose(g.rho, rho, rtol=1e-5)
    P = (gam - 1) * rho * 1.0
    dwdr = (1.0 / np.pi) / h ** 4 * (2.25 * q * q - 3 * q)
    expected_left = 1.0 * 2 * P / rho ** 2 * dwdr
    np.testing.assert_allclose(g.dvdt[0], [expected_left, -expected_left], rtol=1e-4)
    np.testing.assert_allclose(g.dvdt[1:], 0.0, atol=1e-6)
    np.testing.assert_allclose(g.dudt, 0.0, atol=1e-6)


@pytest.mark.parametrize("prec", PRECS)
def test_step_parity(prec):
    cfg, ps = _turb(16, 17)
    nsteps = 3
    with Context([cfg], [ps.n], precision=GPREC[prec]) as ctx:
        ctx.upload(0, ps)
        ctx.prime()
        c0 = ctx.conserved()[0]
        ctx.step(nsteps)
        c1 = ctx.conserved()[0]
        g = ctx.download(0).sorted_by_id()
    o = oracle_ref.OracleRun(cfg, ps, prec)
    oc0 = o.prime()
    assert abs(c0["dt_next"] - oc0["dt_next"]) <= 1e-5 * oc0["dt_next"]
    oc1 = o.step(nsteps)
    order = o.by_id()
    for k in ("mass", "energy"):
        assert abs(c1[k] - oc1[k]) <= RTOL_GLOBAL * abs(oc1[k]), (k, c1[k], oc1[k])
    pscale = float(np.sum(o.mass.astype(np.float64) * np.sqrt(np.sum(o.vel.astype(np.float64) ** 2, axis=0))))
    for a in range(3):
        assert abs(c1["momentum"][a] - oc1["momentum"][a]) <= RTOL_GLOBAL * pscale
    # positions after 3 steps: the trajectories agree to FP32 rounding growth
    dx = np.abs(g.pos - o.pos[:, order])
    dx = np.minimum(dx, 1.0 - dx)
    assert float(np.max(dx)) < 1e-5
    assert float(np.max(np.abs(g.rho - o.rho[order]) / o.rho[order])) < 1e-4


def test_graph_replay_matches_eager():
    cfg, ps = _turb(12, 23)
    outs = []
    for use_graph in (False, True):
        with Context([cfg], [ps.n], precision="mixed") as ctx:
            ctx.upload(0, ps)
            ctx.prime()
            if use_graph:
                ctx.graph_capture()
                ctx.graph_launch(2)
                ctx.sync()
                ctx.graph_launch(3)
            else:
                ctx.step(5)
            ctx.sync()
            outs.append((ctx.download(0).sorted_by_id(), ctx.conserved()[0]))
    (a, ca), (b, cb) = outs
    for f in ("pos", "vel", "u", "rho", "B", "psi", "dvdt"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert ca == cb


def test_graph_launch_of_zero_steps_changes_nothing():
    """A zero-step graph launch after an upload leaves the stage state as the upload left it: the
    force stage still needs a density pass first."""
    cfg, ps = _turb(10, 24)
    with Context([cfg], [ps.n], precision="mixed") as ctx:
        ctx.upload(0, ps)
        ctx.prime()
        ctx.graph_capture()
        ctx.upload(0, ps)
        ctx.graph_launch(0)
        with pytest.raises(capi.SphError, match="density pass"):
            ctx.forces()
        ctx.graph_launch(1)  # the eager odd step, then the state is a stepped one again
        ctx.forces()
        ctx.sync()


def test_graph_dropped_when_the_force_variant_changes():
    """A graph captured while every particle shares one smoothing length holds the one-h force
    variant. Uploading particles of two smoothing lengths drops it (replay would be wrong); a graph
    captured again replays like eager steps."""
    cfg, ps = _turb(10, 25)
    two = ps.copy()
    two.h[1::2] = np.float32(0.9) * two.h[1::2]
    outs = []
    with Context([cfg], [ps.n], precision="mixed") as ctx:
        ctx.upload(0, ps)
        ctx.prime()
        ctx.graph_capture()
        ctx.upload(0, two)
        with pytest.raises(capi.SphError, match="no captured graph"):
            ctx.graph_launch(2)
        ctx.prime()
        ctx.graph_capture()
        ctx.graph_launch(3)
        ctx.sync()
        outs.append(ctx.download(0).sorted_by_id())
    with Context([cfg], [ps.n], precision="mixed") as ctx:
        ctx.upload(0, two)
        ctx.prime()
        ctx.step(3)
        ctx.sync()
        outs.append(ctx.download(0).sorted_by_id())
    for f in ("pos", "vel", "u", "rho", "B", "psi", "dvdt"):
        assert np.array_equal(getattr(outs[0], f), getattr(outs[1], f)), f


def test_batched_runs_equal_alone():
    """Each run of a ragged batch gives bit-identical results to the same run alone."""
    runs = [ic.kelvin_helmholtz(nt, 100 + r) for r, nt in enumerate([4200, 9000, 5100])]
    alone = []
    for cfg, ps in runs:
        with Context([cfg], [ps.n], precision="mixed") as ctx:
            ctx.upload(0, ps)
            ctx.prime()
            ctx.step(2)
            ctx.sync()
            alone.append(ctx.download(0).sorted_by_id())
    with Context([c for c, _ in runs], [p.n for _, p in runs], precision="mixed") as ctx:
        for r, (_, ps) in enumerate(runs):
            ctx.upload(r, ps)
        ctx.prime()
        ctx.step(2)
        ctx.sync()
        mixed = [ctx.download(r).sorted_by_id() for r in range(len(runs))]
    for a, b in zip(alone, mixed):
        for f in ("pos", "vel", "u", "rho", "B", "psi", "dvdt", "dbdt"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_fp64_oracle_error_mixed():
    """Mixed-precision GPU results against the FP64 oracle, within the stated tolerance."""
    cfg, ps = _turb(16, 31)
    g = gpu_stages(cfg, ps, "mixed")
    o = oracle_stages(cfg, ps, "f64")
    order = o.by_id()
    rel_rho = np.abs(g.rho - o.rho[order]) / o.rho[order]
    assert float(np.max(rel_rho)) < 1e-5
    e = np.abs(g.dvdt - o.dvdt[:, order]) / o.sc_a[order]
    assert float(np.max(e)) < 1e-5


# Launch shapes of the pair passes: chunked with per-target lane groups (pair_balance 0), with
# compacted round-robin pairs (1), and whole-cell lists (2). All must give the oracle's results.
MODES = {
    "chunked": dict(density_chunk=128, force_chunk=64, max_lanes_per_target=64, pair_balance=0),
    "balanced": dict(density_chunk=256, force_chunk=128, max_lanes_per_target=16, pair_balance=1),
    "cell512": dict(density_chunk=512, force_chunk=128, max_lanes_per_target=64, pair_balance=2),
    "cell512k16": dict(density_chunk=512, force_chunk=128, max_lanes_per_target=16, pair_balance=2),
    "chunked128": dict(density_chunk=256, force_chunk=128, max_lanes_per_target=16, pair_balance=0),
    # persistent grids of one wave per compute unit: cases with more supercells than waves stride
    "chunked_grid": dict(density_chunk=128, force_chunk=64, max_lanes_per_target=64, pair_balance=0, waves_per_cu=1),
    "balanced_grid": dict(density_chunk=256, force_chunk=128, max_lanes_per_target=16, pair_balance=1,
                          waves_per_cu=1),
    "cell512_grid": dict(density_chunk=512, force_chunk=128, max_lanes_per_target=64, pair_balance=2,
                         waves_per_cu=1),
}


def _clump_case():
    """A sparse periodic lattice (about 2 particles per supercell) with a tight clump of 50
    particles inside one supercell: that cell has at most 64 targets and its masks are stored, but
    its accepted pairs (> 2450) exceed the whole-cell list, so it takes the direct per-lane path."""
    n, nclump = 10, 50
    dx = 1.0 / n
    rng = np.random.default_rng(41)
    lat = ic.lattice(n, n, n, (0, 0, 0), dx, rng)
    centre = np.array([[0.52], [0.47], [0.55]])
    pos = np.concatenate([lat, centre + rng.uniform(-0.01, 0.01, size=(3, nclump))], axis=1)
    ps = ParticleSet(pos.shape[1])
    ps.id = np.arange(ps.n, dtype=np.uint64)
    ps.pos = pos
    h0 = 0.06
    ps.mass = np.full(ps.n, dx ** 3, dtype=np.float32)
    ps.h = np.full(ps.n, h0, dtype=np.float32)
    ps.u = np.full(ps.n, 1.5, dtype=np.float32)
    ps.rho = np.ones(ps.n, dtype=np.float32)
    ps.vel = rng.normal(0, 0.1, size=(3, ps.n)).astype(np.float32)
    ps.B = np.stack([np.full(ps.n, 0.1), rng.normal(0, 0.02, ps.n), np.zeros(ps.n)]).astype(np.float32)
    cfg = capi.make_config(periodic=(1, 1, 1), kernel=capi.KERNEL_WENDLAND, h0=h0, mhd=1, cleaning=1, borve=1)
    return cfg, ps


def _big_cells_case():
    """Supercells of edge >= 2H on a 16^3 lattice: ~150 targets (several target groups) and ~4000
    candidates per cell (several chunks of the staging buffer; masks not kept, so the force pass
    culls again)."""
    cfg, ps = ic.mhd_turbulence(n=16, seed=43)
    cfg.cell_factor = 2.0
    return cfg, ps


def _crowded_case():
    """A periodic 20^3 lattice in 7^3 supercells (~23 particles each) with 12 and 6 extra particles
    spread through two of them. Most supercells hand the force pass their kept candidate list in the
    mask row; the one at ~29 particles (2 lanes per target) keeps too many candidates for the row, so
    the force pass recomputes the list; the one at ~35 (1 lane per target) has more kept candidates
    than its masks hold, so the force pass culls again with the recomputed list."""
    n, h0 = 20, 0.065
    dx = 1.0 / n
    rng = np.random.default_rng(53)
    lat = ic.lattice(n, n, n, (0, 0, 0), dx, rng)
    e = 1.0 / 7
    a = np.array([[2.0], [3.0], [4.0]]) * e
    b = np.array([[5.0], [1.0], [2.0]]) * e
    pos = np.concatenate([lat, a + rng.uniform(0.05 * e, 0.95 * e, size=(3, 12)),
                          b + rng.uniform(0.05 * e, 0.95 * e, size=(3, 6))], axis=1)
    ps = ParticleSet(pos.shape[1])
    ps.id = np.arange(ps.n, dtype=np.uint64)
    ps.pos = pos
    ps.mass = np.full(ps.n, dx ** 3, dtype=np.float32)
    ps.h = np.full(ps.n, h0, dtype=np.float32)
    ps.u = np.full(ps.n, 1.5, dtype=np.float32)
    ps.rho = np.ones(ps.n, dtype=np.float32)
    ps.vel = rng.normal(0, 0.1, size=(3, ps.n)).astype(np.float32)
    ps.B = np.stack([np.full(ps.n, 0.1), rng.normal(0, 0.02, ps.n), np.zeros(ps.n)]).astype(np.float32)
    cfg = capi.make_config(periodic=(1, 1, 1), kernel=capi.KERNEL_WENDLAND, h0=h0, mhd=1, cleaning=1, borve=1)
    return cfg, ps


def _two_h_case():
    """MHD turbulence with every other particle at 0.9 h: pairs of unequal smoothing lengths take the
    second kernel gradient and the mean h, in the general (not one-h) instantiation of each shape."""
    cfg, ps = _turb(14, 19)
    ps.h[1::2] = np.float32(0.9) * ps.h[1::2]
    return cfg, ps


CASES = {
    "turbulence": lambda: _turb(16, 17),
    "two_h": _two_h_case,
    "sod_walls": lambda: ic.sod_tube(seed=5, nx_left=32, cross=4),
    "clump": _clump_case,
    "crowded": _crowded_case,
    "big_cells": _big_cells_case,
}


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("prec", PRECS)
def test_forces_parity_every_launch_shape(mode, case, prec):
    cfg, ps = CASES[case]()
    print(f"shape {mode} case {case} prec {prec} n={ps.n}", flush=True)
    g = gpu_stages(cfg, ps, GPREC[prec], tuning=MODES[mode])
    o = oracle_stages(cfg, ps, prec)
    compare_forces(g, o, bool(cfg.mhd))


def test_mixed_kernel_batch():
    """Runs with different smoothing kernels in one context take the launch instantiation that reads
    the kernel per run: each run meets the oracle contract, and equals the same run alone."""
    runs = [ic.mhd_turbulence(n=10, seed=31 + k, kernel=kern)
            for k, kern in enumerate([capi.KERNEL_CUBIC, capi.KERNEL_QUINTIC, capi.KERNEL_WENDLAND])]
    with Context([c for c, _ in runs], [p.n for _, p in runs], precision="mixed") as ctx:
        for r, (_, p) in enumerate(runs):
            ctx.upload(r, p)
        ctx.resort()
        ctx.density()
        ctx.forces()
        ctx.sync()
        mixed = [ctx.download(r).sorted_by_id() for r in range(len(runs))]
    for (cfg, ps), g in zip(runs, mixed):
        compare_forces(g, oracle_stages(cfg, ps, "mixed"), mhd=True)
        alone = gpu_stages(cfg, ps, "mixed")
        for f in ("rho", "nnb", "dvdt", "dudt", "dbdt", "divb", "dpsidt", "vsig"):
            assert np.array_equal(getattr(g, f), getattr(alone, f)), (cfg.kernel, f)


def test_one_h_instantiation_equals_general():
    """A run whose particles share one smoothing length takes the force pass's one-h instantiation
    (no test of h_j against h_i, the mean h taken as h). In a context that also holds a run of two
    smoothing lengths the same run takes the general instantiation: its results are bit-identical
    either way, and the two-h run (every other particle at 0.9 h, so pairs of unequal h take the
    second kernel gradient and the mean h) meets the oracle contract."""
    cfg, ps = _turb(12, 5)
    tcfg, tps = _turb(10, 6)
    tps.h[1::2] = np.float32(0.9) * tps.h[1::2]
    assert np.unique(ps.h).size == 1 and np.unique(tps.h).size == 2
    print(f"one-h vs general: n={ps.n}, two-h run n={tps.n}", flush=True)
    alone = gpu_stages(cfg, ps, "mixed")
    with Context([cfg, tcfg], [ps.n, tps.n], precision="mixed") as ctx:
        ctx.upload(0, ps)
        ctx.upload(1, tps)
        ctx.resort()
        ctx.density()
        ctx.forces()
        ctx.sync()
        both = ctx.download(0).sorted_by_id()
        two = ctx.download(1).sorted_by_id()
    for f in ("rho", "nnb", "dvdt", "dudt", "dbdt", "divb", "dpsidt", "vsig"):
        assert np.array_equal(getattr(alone, f), getattr(both, f)), f
    compare_forces(two, oracle_stages(tcfg, tps, "mixed"), mhd=True)


def test_hydro_run_batched_with_mhd_runs_equals_alone():
    """A hydro run in a context that carries MHD fields (another run is MHD) takes the MHD launch
    instantiation with its MHD terms off; its results equal the same run alone in a hydro context."""
    runs = [ic.kelvin_helmholtz(4000, 71), ic.sod_tube(seed=7, nx_left=32, cross=4)]
    got, alone = [], []
    with Context([c for c, _ in runs], [p.n for _, p in runs], precision="mixed") as ctx:
        for r, (_, ps) in enumerate(runs):
            ctx.upload(r, ps)
        ctx.step(3)
        ctx.sync()
        got = [ctx.download(r).sorted_by_id() for r in range(len(runs))]
    for cfg, ps in runs:
        with Context([cfg], [ps.n], precision="mixed") as ctx:
            ctx.upload(0, ps)
            ctx.step(3)
            ctx.sync()
            alone.append(ctx.download(0).sorted_by_id())
    for a, b in zip(alone, got):
        for f in ("pos", "vel", "u", "rho", "dvdt", "dudt", "vsig", "nnb"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("wpc", [1, 4, 16])
def test_persistent_grid_batch_matches_oracle_and_alone(wpc):
    """A ragged batch of Kelvin-Helmholtz runs (about 6,800 supercells) stepped with a persistent
    pair-pass grid of wpc waves per compute unit, each wave striding over supercells: every run
    meets the oracle contract at the first pass, and after two steps the batch is bit-identical to
    the same batch with one wave per supercell and to each run alone."""
    sizes = [4096, 9000, 20000, 6000, 12000, 30000, 5000, 8000]
    runs = [ic.kelvin_helmholtz(nt, 900 + r) for r, nt in enumerate(sizes)]
    cfgs, n = [c for c, _ in runs], [p.n for _, p in runs]
    tune = dict(density_chunk=128, force_chunk=64, max_lanes_per_target=64, pair_balance=0)

    def batch(waves, steps):
        with Context(cfgs, n, precision="mixed") as ctx:
            ctx.set_tuning(**tune, waves_per_cu=waves)
            for r, (_, p) in enumerate(runs):
                ctx.upload(r, p)
            if steps:
                ctx.prime()
                ctx.step(steps)
            else:
                ctx.resort()
                ctx.density()
                ctx.forces()
            ctx.sync()
            return [ctx.download(r).sorted_by_id() for r in range(len(runs))]

    print(f"persistent grid wpc={wpc} runs={len(runs)} particles={sum(n)}", flush=True)
    first = batch(wpc, 0)
    for r in (0, 2, 5):
        compare_forces(first[r], oracle_stages(*runs[r], "mixed"), mhd=True)
    stepped = batch(wpc, 2)
    for a, b in zip(batch(0, 2), stepped):
        for f in ("pos", "vel", "u", "rho", "B", "psi", "dvdt", "dbdt", "nnb"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
    for r in (1, 5):
        cfg, ps = runs[r]
        with Context([cfg], [ps.n], precision="mixed") as ctx:
            ctx.set_tuning(**tune, waves_per_cu=wpc)
            ctx.upload(0, ps)
            ctx.prime()
            ctx.step(2)
            ctx.sync()
            alone = ctx.download(0).sorted_by_id()
        for f in ("pos", "vel", "u", "rho", "B", "dvdt", "nnb"):
            assert np.array_equal(getattr(alone, f), getattr(stepped[r], f)), (r, f)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [dict(density_chunk=128, force_chunk=64, waves_per_cu=0),
                                   dict(density_chunk=256, force_chunk=128, waves_per_cu=0),
                                   dict(density_chunk=128, force_chunk=64, waves_per_cu=1)])
def test_half_h_supercells_parity(prec, shape):
    """Supercells of edge >= H/2 (cell_factor 0.5, the 125-cell stencil): neighbour counts bit-exact and
    density / forces within tolerance against the oracle on the same grid."""
    for name, (cfg, ps) in (("turbulence", _turb(16, 19)), ("sod_walls", ic.sod_tube(seed=6, nx_left=32, cross=4))):
        cfg.cell_factor = 0.5
        print(f"half-H {name} prec {prec} {shape} n={ps.n}", flush=True)
        g = gpu_stages(cfg, ps, GPREC[prec], tuning=dict(shape, max_lanes_per_target=64, pair_balance=0))
        o = oracle_stages(cfg, ps, prec)
        assert oracle_ref.grid(cfg).reach == 2
        compare_forces(g, o, bool(cfg.mhd))


def test_half_h_and_full_h_runs_batched_equal_alone():
    """A context mixing a run on supercells of edge >= H/2 with runs on edge >= H takes the 125-cell
    kernels for all; each run still equals itself alone, bit for bit, after three steps, and the
    whole-cell or balanced pair loops are refused while a finer grid is present."""
    fine = ic.kelvin_helmholtz(5000, 81)
    fine[0].cell_factor = 0.5
    runs = [fine, ic.kelvin_helmholtz(6000, 82), ic.mhd_turbulence(n=14, seed=83)]
    cfgs, n = [c for c, _ in runs], [p.n for _, p in runs]
    with Context(cfgs, n, precision="mixed") as ctx:
        with pytest.raises(capi.SphError):
            ctx.set_tuning(pair_balance=1)
        for r, (_, p) in enumerate(runs):
            ctx.upload(r, p)
        ctx.prime()
        ctx.step(3)
        ctx.sync()
        got = [ctx.download(r).sorted_by_id() for r in range(len(runs))]
    for (cfg, ps), g in zip(runs, got):
        with Context([cfg], [ps.n], precision="mixed") as ctx:
            ctx.upload(0, ps)
            ctx.prime()
            ctx.step(3)
            ctx.sync()
            a = ctx.download(0).sorted_by_id()
        for f in ("pos", "vel", "u", "rho", "B", "dvdt", "dbdt", "nnb"):
            assert np.array_equal(getattr(a, f), getattr(g, f)), (cfg.cell_factor, f)
