"""GPU: the serving pool (sph_pool_create / sph_pool_set_runs) and the continuous-batching server.

Contract: a run's results are bit-identical whether it runs alone in its own context, in a pool
whose run set changes around it (kept runs move to new offsets, new runs join and are primed), or
under the server's admission and retirement."""
from __future__ import annotations

import numpy as np
import pytest

from spearhed_amd import ic, setups
from spearhed_amd.serving import Request, Server, run_bytes
from spearhed_amd.stepper import Context, Pool

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "u", "rho", "B", "psi", "dvdt", "dbdt", "nnb")


def _alone(cfg, ps, steps, blocks=None, seed=0):
    n = ps.n if ps is not None else Request(0, cfg, None, 0, blocks=blocks).n
    with Context([cfg], [n], precision="mixed") as ctx:
        if ps is not None:
            ctx.upload(0, ps)
        else:
            ctx.init_run(0, blocks, seed=seed)
        ctx.prime()
        if steps:
            ctx.step(steps)
        ctx.sync()
        return ctx.download(0).sorted_by_id(), ctx.conserved()[0]


def _same(a, b):
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_pool_runs_move_and_join():
    runs = [ic.kelvin_helmholtz(nt, 500 + r) for r, nt in enumerate([3000, 6000, 4100])]
    late = ic.orszag_tang_n(3500, 7)
    print("pool: runs", [p.n for _, p in runs], "late", late[1].n, flush=True)
    with Pool(40000, 4000, 8, "mixed", True) as pool:
        pool.set_runs([], [(c, p.n) for c, p in runs])
        for r, (_, p) in enumerate(runs):
            pool.upload(r, p)
        pool.prime()
        pool.step(2)
        # drop run 1, reverse the others, add one
        new = pool.set_runs([2, 0], [(late[0], late[1].n)])
        assert new == [2]
        pool.upload(2, late[1])
        pool.prime()
        pool.step(3)
        pool.sync()
        got = [pool.download(r).sorted_by_id() for r in range(3)]
        cons = pool.conserved()
    for (cfg, ps), steps, g, c in zip([runs[2], runs[0], late], [5, 5, 3], got, cons):
        a, ca = _alone(cfg, ps, steps)
        _same(a, g)
        assert ca["time"] == c["time"] and ca["dt_next"] == c["dt_next"]


def test_kept_run_not_yet_stepped_still_primes():
    """A run uploaded and then kept by the next set_runs before any step still takes its first step
    as the priming step (with a fixed dt a lost flag would drift it instead)."""
    (cfg_a, ps_a), (cfg_b, ps_b) = ic.kelvin_helmholtz(3000, 61), ic.kelvin_helmholtz(4000, 62)
    cfg_a.dt_fixed = cfg_b.dt_fixed = 1e-4
    with Pool(20000, 4000, 4, "mixed", True) as pool:
        pool.set_runs([], [(cfg_a, ps_a.n)])
        pool.upload(0, ps_a)
        assert pool.set_runs([0], [(cfg_b, ps_b.n)]) == [1]
        pool.upload(1, ps_b)
        pool.step(3)  # the priming step and two steps, for both runs
        pool.sync()
        got = [pool.download(r).sorted_by_id() for r in range(2)]
        cons = pool.conserved()
    for (cfg, ps), g, c in zip([(cfg_a, ps_a), (cfg_b, ps_b)], got, cons):
        a, ca = _alone(cfg, ps, 2)
        _same(a, g)
        assert ca["time"] == c["time"] == pytest.approx(2e-4, rel=1e-12)


def test_pool_rejects_over_capacity():
    cfg, ps = ic.kelvin_helmholtz(3000, 1)
    with Pool(ps.n + 10, 4000, 4, "mixed", True) as pool:
        pool.set_runs([], [(cfg, ps.n)])
        with pytest.raises(Exception):
            pool.set_runs([0], [(cfg, ps.n)])
        pool.upload(0, ps)  # the failed call left the pool as it was
        pool.prime()
        pool.step(1)
        pool.sync()
        g = pool.download(0).sorted_by_id()
    _same(_alone(cfg, ps, 1)[0], g)


def test_served_results_equal_alone():
    probs = [ic.kelvin_helmholtz(2500, 40), ic.brio_wu(2000, 41), ic.rotor(3000, 42),
             ic.kelvin_helmholtz(5000, 43), ic.orszag_tang_n(2500, 44)]
    steps = [3, 5, 2, 4, 3, 4, 2]
    reqs = [Request(r, c, p, steps[r], arrival=float(r)) for r, (c, p) in enumerate(probs)]
    # two requests filled on the device from lattice blocks
    for r, make in ((5, setups.rotor), (6, setups.brio_wu)):
        cfg, blocks, _ = make(3000)
        reqs.append(Request(r, cfg, None, steps[r], arrival=float(r), blocks=blocks, seed=100 + r))
    budget = 3 * max(run_bytes(q.cfg, q.n) for q in reqs)  # forces retire/admit cycles
    srv = Server(budget, precision="mixed", max_runs=8)
    try:
        summary = srv.serve(reqs, realtime=False)
    finally:
        srv.close()
    print("serve:", summary, flush=True)
    assert sorted(srv.results) == list(range(len(reqs)))
    for q in reqs:
        a, ca = _alone(q.cfg, q.ps, q.steps, q.blocks, q.seed)
        res = srv.results[q.rid]
        _same(a, res.ps.sorted_by_id())
        assert res.conserved["time"] == ca["time"]


def test_ensemble_autotuned_graph_matches_alone():
    """The benchmark's ensemble sequence (prime, autotune, prime, graph capture and replay) on a
    ragged batch gives each run's physics: neighbour counts as alone, density and positions within
    the FP32 summation-order tolerance of the run stepped alone with default launch shapes."""
    sizes = [4096, 9000, 20000, 5200, 13000, 6100, 4500, 16000, 7000, 11000, 4200, 8000]
    runs = [ic.kelvin_helmholtz(nt, 900 + r) for r, nt in enumerate(sizes)]
    steps = 4
    with Context([c for c, _ in runs], [p.n for _, p in runs], precision="mixed") as ctx:
        for r, (_, p) in enumerate(runs):
            ctx.upload(r, p)
        ctx.prime()
        tuned = ctx.autotune(2)
        ctx.prime()
        ctx.graph_capture()
        ctx.graph_launch(steps)
        ctx.sync()
        got = [ctx.download(r).sorted_by_id() for r in range(len(runs))]
        cons = ctx.conserved()
    print("tuned:", tuned, flush=True)
    for (cfg, ps), g, c in zip(runs, got, cons):
        a, ca = _alone(cfg, ps, steps)
        assert np.isfinite(c["energy"]) and c["time"] > 0
        assert abs(c["time"] - ca["time"]) <= 1e-5 * ca["time"]
        assert abs(np.mean(g.nnb) - np.mean(a.nnb)) <= 0.01 * np.mean(a.nnb)
        assert np.max(np.abs(g.rho - a.rho) / a.rho) <= 1e-4
        assert np.max(np.abs(g.pos - a.pos)) <= 1e-6


def test_many_runs_keep_their_tables():
    """A batch of 64 runs in one arena: every run's uploaded state reads back unchanged, and after
    the first prime every run has its neighbours. (The arena clear at creation must be ordered
    before the run tables are uploaded on the context's stream.)"""
    runs = [ic.kelvin_helmholtz(4100 + 37 * r, 1300 + r) for r in range(64)]
    with Context([c for c, _ in runs], [p.n for _, p in runs], precision="mixed") as ctx:
        for r, (_, p) in enumerate(runs):
            ctx.upload(r, p)
        for r in (0, 1, 31, 63):
            d = ctx.download(r)
            assert np.array_equal(d.pos, runs[r][1].pos) and np.array_equal(d.h, runs[r][1].h), r
        ctx.prime()
        ctx.sync()
        means = [ctx.neighbour_total(r) / p.n for r, (_, p) in enumerate(runs)]
    assert min(means) > 40, means


@pytest.mark.parametrize("dt_fixed", [0.0, 2e-4])
def test_first_step_after_upload_primes(dt_fixed):
    """After new initial data the run's first step is its priming step (no drift, no kick, no time
    advance; first CFL dt), so upload + (k+1) steps equals upload + prime + k steps bit for bit."""
    cfg, ps = ic.kelvin_helmholtz(5000, 77)
    cfg.dt_fixed = dt_fixed
    outs = []
    for primed in (True, False):
        with Context([cfg], [ps.n], precision="mixed") as ctx:
            ctx.upload(0, ps)
            if primed:
                ctx.prime()
                ctx.step(3)
            else:
                ctx.step(4)
            ctx.sync()
            outs.append((ctx.download(0).sorted_by_id(), ctx.conserved()[0]))
    (a, ca), (b, cb) = outs
    _same(a, b)
    assert ca == cb


def test_served_response_fields_equal_alone():
    """The asynchronous copy-out path (pinned ring, sph_download_async) returns the same selected
    fields and conserved record as the run stepped alone."""
    probs = [ic.kelvin_helmholtz(3000, 60), ic.orszag_tang_n(2500, 61), ic.rotor(2600, 62), ic.brio_wu(2400, 63)]
    steps = [3, 2, 4, 3]
    reqs = [Request(r, c, p, steps[r], arrival=float(r)) for r, (c, p) in enumerate(probs)]
    fields = ("id", "pos", "vel", "rho", "u", "B", "psi", "nnb")
    budget = 2 * max(run_bytes(q.cfg, q.n) for q in reqs)
    srv = Server(budget, precision="mixed", max_runs=8, response_fields=fields)
    try:
        srv.serve(reqs, realtime=False)
    finally:
        srv.close()
    for q in reqs:
        a, ca = _alone(q.cfg, q.ps, q.steps)
        got = srv.results[q.rid]
        assert got.ps is None
        order = np.argsort(got.fields["id"])
        assert np.array_equal(got.fields["id"][order], a.id)
        for f in ("pos", "vel", "rho", "u", "B", "psi", "nnb"):
            assert np.array_equal(got.fields[f][..., order], getattr(a, f)), f
        for k in ("mass", "energy", "time", "dt_next"):
            assert got.conserved[k] == ca[k], k


def test_pool_copy_out_holds_the_run_at_the_call():
    """In a pool, sph_download_async snapshots the run into the pool's device ring on the pool's
    stream and copies the snapshot out on a second stream while later steps run. Each copy holds the
    run as it was where the call was enqueued, though the steps after it rewrite the run's state;
    consecutive copy-outs take the ring's space back (the pool's stream waits for the older copies
    out); sph_mark_wait on a later mark sees every copy landed."""
    from spearhed_amd.stepper import PinnedRing, download_async
    runs = [ic.kelvin_helmholtz(nt, 700 + r) for r, nt in enumerate([3000, 5000])]
    fields = ("id", "pos", "vel", "rho", "u", "B", "psi", "nnb")
    # the ring holds 16 B per particle of capacity (384 kB here): one response of these fields for
    # the 5000-particle run (~363 kB) fits, two responses in a row do not
    cap = 24000
    print("pool copy-out: runs", [p.n for _, p in runs], flush=True)
    rings = [PinnedRing(1 << 20) for _ in range(3)]
    taken = []  # (run, steps taken when copied, arrays, conserved)
    try:
        with Pool(cap, 4000, 4, "mixed", True) as pool:
            pool.set_runs([], [(c, p.n) for c, p in runs])
            for r, (_, p) in enumerate(runs):
                pool.upload(r, p)
            pool.prime()
            pool.step(2)
            plan = [(1, 2), (0, 3), (1, 4)]  # run copied out after that many steps
            for k, (run, steps) in enumerate(plan):
                arrays, cons = download_async(pool, run, fields, rings[k])
                taken.append((run, steps, arrays, cons))
                pool.step(1)  # rewrites the state the copy was taken from
            mark = pool.mark()
            pool.step(2)
            pool.wait_mark(mark)
            snap = [(run, steps, {f: np.array(a) for f, a in arrays.items()}, cons.as_dict())
                    for run, steps, arrays, cons in taken]
            pool.sync()
            final = [pool.download(r).sorted_by_id() for r in range(2)]
    finally:
        for ring in rings:
            ring.close()
    for run, steps, got, cons in snap:
        cfg, ps = runs[run]
        a, ca = _alone(cfg, ps, steps)
        order = np.argsort(got["id"])
        assert np.array_equal(got["id"][order], a.id)
        for f in ("pos", "vel", "rho", "u", "B", "psi", "nnb"):
            assert np.array_equal(got[f][..., order], getattr(a, f)), (run, steps, f)
        for k in ("time", "dt_next", "energy"):
            assert cons[k] == ca[k], (run, steps, k)
    for r, (cfg, ps) in enumerate(runs):
        _same(_alone(cfg, ps, 7)[0], final[r])


def test_set_runs_rejects_block_overflow_and_leaves_runs_intact():
    """A context from sph_ctx_create has exactly the particle-kernel blocks of its runs: two runs of
    256 particles hold 2 blocks, so a new set of 1 + 511 particles (3 blocks) fits the particle count
    but not the blocks. The call must fail before anything moves and leave both runs as they were."""
    import ctypes as C

    from spearhed_amd import capi

    cfg, ps = ic.kelvin_helmholtz(256, 71)
    ps = ps.take(np.arange(256))
    ps.id = np.arange(256, dtype=np.uint64)
    with Context([cfg, cfg], [256, 256], precision="mixed") as ctx:
        ctx.upload(0, ps)
        ctx.upload(1, ps)
        ctx.prime()
        ctx.step(1)
        ctx.sync()
        before = [ctx.download(r).sorted_by_id() for r in range(2)]
        add = (capi.RunConfig * 2)(cfg, cfg)
        sizes = (C.c_int64 * 2)(1, 511)
        rc = ctx.lib.sph_pool_set_runs(ctx.h, None, 0, add, sizes, 2)
        assert rc == capi.SPH_E_CAPACITY
        after = [ctx.download(r).sorted_by_id() for r in range(2)]
        for a, b in zip(before, after):
            _same(a, b)
        ctx.step(1)  # still steps
        ctx.sync()


def test_empty_pool_stages_are_no_ops():
    with Pool(10000, 1000, 4, "mixed", True) as pool:
        for stage in ("drift", "resort", "density", "forces"):
            getattr(pool, stage)()
        pool.integrate()
        pool.step(2)
        assert pool.conserved() == []
        pool.sync()
