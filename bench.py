#!/usr/bin/env python3
"""Benchmark: SPH-MHD particle-steps/s at 2M particles on one MI355X, plus ensemble step latency.

Workload (the metric's configuration): mixed-precision 3D MHD turbulence in a periodic unit cube,
128^3 = 2,097,152 particles (jittered lattice, rho = P = 1, gamma = 5/3, Gaussian random velocity
with E(k) ~ k^-2 for k in [1, 8] at rms Mach 1, B = 0.1 z + 10% divergence-free perturbation),
Wendland C2 with ~57 neighbours, Monaghan viscosity, SPMHD with Borve correction and hyperbolic
cleaning, CFL dt. A step = drift + resort + density + forces + kick with FP64 CFL/conservation
reductions, all on the GPU with inputs resident in HBM.

Prints one JSON line. value = particles * K / wall time of K steps (bracketed by device syncs).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from spearhed_amd import ic  # noqa: E402
from spearhed_amd.stepper import Context  # noqa: E402

FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (MFMA unused: pair physics is not a dense contraction)
HBM_PEAK_GBS = 8000.0

# Algorithmic FP32 work per accepted pair and per particle (+ - * / sqrt compare = 1, FMA = 2),
# measured by the operation-counting build of the oracle on seeded turbulence runs and fitted as
# a * pairs + b * particles + c * candidates (tests/test_flop_count.py keeps these equal to the fit;
# the candidate term, work on rejected candidates, is not counted). The force count includes the
# viscosity branch at this workload's share of approaching pairs and counts 1/sqrt as two operations.
FORCE_FLOP_PER_PAIR = 108
FORCE_FLOP_PER_PARTICLE = 80
DENSITY_FLOP_PER_PAIR = 15
# Algorithmic HBM bytes of the force pass per particle: read the 5 float4 interaction records once,
# write dv/dt (12), du/dt, dB/dt (12), dpsi/dt, div B, v_sig (16).
FORCE_BYTES_PER_PARTICLE = 80 + 44
DENSITY_FLOP_PER_PARTICLE = 49
# Algorithmic bytes per particle of the other stages of a mixed-precision MHD step (each field read and
# written once): drift + key + histogram rank, scatter, gather into the sorted set (+ pair record),
# density (positions, masses, h in; rho, counts and the force pass's records out), kick + sums.
STAGE_BYTES_PER_PARTICLE = {"drift_bin": 70, "resort": 32 + 190, "density": 130, "forces": FORCE_BYTES_PER_PARTICLE,
                            "integrate": 110}
CONFIG_STEPS = 100  # the 2M-particle turbulence configuration's step count


def measured_traffic(kernel: str, particles: int, tuning: dict):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    (each scaled by the copy-kernel calibration), when they were taken on this workload and tuning."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        rec = json.load(open(path))[kernel]
    except (OSError, KeyError, ValueError):
        return None
    if rec.get("particles") != particles or rec.get("tuning") != tuning:
        return None
    return rec["bytes_per_launch"]


def step_roofline(n: int, pairs: int, step_ms: float) -> dict:
    """The whole step against its speed of light: each stage bounded by the larger of its FP32 work at
    the FP32 peak and its algorithmic bytes at the HBM peak, summed over the stages."""
    flop = {"density": DENSITY_FLOP_PER_PAIR * pairs + DENSITY_FLOP_PER_PARTICLE * n,
            "forces": FORCE_FLOP_PER_PAIR * pairs + FORCE_FLOP_PER_PARTICLE * n}
    sol = {}
    for stage, b in STAGE_BYTES_PER_PARTICLE.items():
        sol[stage] = max(flop.get(stage, 0) / (FP32_PEAK_TFLOPS * 1e12), b * n / (HBM_PEAK_GBS * 1e9)) * 1e3
    total = sum(sol.values())
    return {"sol_ms": total, "stage_sol_ms": sol, "measured_ms": step_ms, "frac": total / step_ms,
            "basis": "sum over stages of max(FP32 FLOP / 157.3 TFLOP/s, algorithmic bytes / 8 TB/s)"}


def run_main(args):
    t0 = time.perf_counter()
    cfg, ps = ic.mhd_turbulence(n=args.n, seed=3)
    if args.cell_factor:
        cfg.cell_factor = args.cell_factor
    n = ps.n
    t_ic = time.perf_counter() - t0
    ctx = Context([cfg], [n], precision=args.precision)
    if args.tuning:
        ctx.set_tuning(**json.loads(args.tuning))
    ctx.upload(0, ps)
    if args.tune:
        ctx.prime()
        ctx.autotune(2)
    ctx.prime()
    primed = None if args.no_cpu else ctx.download(0).sorted_by_id()  # density + forces of the initial state
    for _ in range(args.warmup):
        ctx.step(1)
    ctx.sync()
    ctx.profile_begin(args.steps)
    t1 = time.perf_counter()
    ctx.step(args.steps)
    ctx.sync()
    t2 = time.perf_counter()
    prof = ctx.profile_end()
    wall = t2 - t1
    pairs = ctx.neighbour_total(0)
    cons = ctx.conserved()[0]
    tuning = ctx.tuning
    # the configuration's whole run: the same initial state primed and stepped CONFIG_STEPS times (the
    # pair passes slow down as the turbulent state develops, so the first steps flatter the run)
    ctx.upload(0, ps)
    ctx.prime()
    ctx.sync()
    ctx.profile_begin(CONFIG_STEPS)
    t3 = time.perf_counter()
    ctx.step(CONFIG_STEPS)
    ctx.sync()
    t4 = time.perf_counter()
    full = ctx.profile_end()
    full["wall_s"] = t4 - t3
    full["pairs_last"] = ctx.neighbour_total(0)
    ctx.close()
    return dict(n=n, wall=wall, prof=prof, pairs=pairs, cons=cons, t_ic=t_ic, cfg=cfg, ps=ps, tuning=tuning,
                primed=primed, full=full)


def run_ensemble(args):
    """Ragged ensemble of Kelvin-Helmholtz runs: one launch per stage for all runs, captured in a
    hipGraph after the autotuner picked the launch shape, against the same runs stepped one at a
    time in plain FP32 with default launch shapes and no graph."""
    runs = ic.kh_ensemble(args.ensemble_runs, seed=4)
    sizes = [p.n for _, p in runs]
    total = sum(sizes)
    k = args.ensemble_steps
    ctx = Context([c for c, _ in runs], sizes, precision="mixed")
    for r, (_, p) in enumerate(runs):
        ctx.upload(r, p)
    ctx.prime()
    tuning = ctx.autotune(2) if not args.no_tune_ensemble else ctx.tuning
    ctx.prime()
    ctx.graph_capture()
    ctx.graph_launch(2)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.graph_launch(k)
    ctx.sync()
    wall_graph = time.perf_counter() - t1
    ctx.profile_begin(k)
    ctx.step(k)
    ctx.sync()
    prof = ctx.profile_end()
    nb_mean = sum(ctx.neighbour_total(r) for r in range(len(runs))) / total
    energy = [c["energy"] for c in ctx.conserved()]
    ctx.close()
    if not (all(math.isfinite(e) for e in energy) and nb_mean > 10):
        raise RuntimeError(f"ensemble state is not physical (mean neighbours {nb_mean})")
    # plain FP32, one run at a time (contexts created and primed outside the timed regions)
    k_plain = max(2, k // 2)
    wall_plain = 0.0
    for cfg, p in runs:
        with Context([cfg], [p.n], precision="fp32") as c1:
            c1.upload(0, p)
            c1.prime()
            c1.sync()
            t1 = time.perf_counter()
            c1.step(k_plain)
            c1.sync()
            wall_plain += time.perf_counter() - t1
    batched = total * k / wall_graph
    plain = total * k_plain / wall_plain
    return {"runs": len(runs), "particles": total, "steps": k, "neighbours_mean": nb_mean,
            "particle_steps_per_s": batched,
            "p50_ms": prof["step_ms_p50"], "p99_ms": prof["step_ms_p99"], "tuning": tuning,
            "plain_fp32_one_run_at_a_time": {"particle_steps_per_s": plain, "steps_per_run": k_plain},
            "speedup_vs_plain": batched / plain}


def run_serving(args):
    """Continuous batching of a Poisson stream of independent runs (Brio-Wu, rotor, Orszag-Tang drawn
    uniformly; sizes log-uniform in [8k, 1M]; 20 steps each), admitted by HBM budget. Initial
    conditions are generated before the clock starts; uploads and priming of admitted runs happen
    between batched steps and are reported separately from the step latency."""
    from spearhed_amd.serving import Request, Server, poisson_arrivals

    from spearhed_amd import setups

    rng = np.random.default_rng(args.serve_seed)
    names = sorted(setups.DEVICE_SERVING_PROBLEMS)
    arrivals = poisson_arrivals(args.serve_requests, args.serve_rate, seed=args.serve_seed)
    reqs = []
    for r in range(args.serve_requests):
        n = int(round(math.exp(rng.uniform(math.log(8_000), math.log(1_000_000)))))
        name = names[int(rng.integers(len(names)))]
        if args.serve_host_ic:  # particles generated on the host and uploaded
            cfg, ps = ic.SERVING_PROBLEMS[name](n, args.serve_seed * 1000 + r)
            reqs.append(Request(r, cfg, ps, 20, arrival=float(arrivals[r])))
        else:  # lattice blocks filled on the device
            cfg, blocks, _ = setups.DEVICE_SERVING_PROBLEMS[name](n)
            reqs.append(Request(r, cfg, None, 20, arrival=float(arrivals[r]), blocks=blocks,
                                seed=args.serve_seed * 1000 + r))
    srv = Server(int(args.serve_budget_gb * 2 ** 30), precision="mixed", max_runs=256,
                 response_fields=("id", "pos", "vel", "rho", "u", "B"))
    try:
        t0 = time.perf_counter()
        summary = srv.serve(reqs, realtime=True)
        wall = time.perf_counter() - t0
        lat = np.array([res.finished - res.arrival for res in srv.results.values()]) * 1e3
    finally:
        srv.close()
    work = sum(q.n * q.steps for q in reqs)
    return {"requests": len(reqs), "rate_per_s": args.serve_rate, "hbm_budget_GB": args.serve_budget_gb,
            "pool_reserved_GB": srv.pool_bytes / 2 ** 30, "pool_capacity_particles": srv.max_particles,
            "initial_conditions": "host upload" if args.serve_host_ic else "device lattice blocks",
            "response": "id, pos, vel, rho, u, B + FP64 sums, copied out asynchronously through pinned memory",
            "particles_total": int(sum(q.n for q in reqs)), "step_ms_p50": summary["step_ms_p50"],
            "step_ms_p99": summary["step_ms_p99"], "step_wall_ms_p50": summary["step_wall_ms_p50"],
            "step_wall_ms_p99": summary["step_wall_ms_p99"], "mean_batch_runs": summary["mean_batch_runs"],
            "particle_steps_per_s_in_steps": summary["particle_steps_per_s"],
            "particle_steps_per_s_wall": work / wall, "request_latency_ms_p50": float(np.percentile(lat, 50)),
            "request_latency_ms_p99": float(np.percentile(lat, 99)),
            "admission_and_upload_ms_total": summary["admission_ms_total"],
            "retire_ms_total": summary["retire_ms_total"], "move_ms_total": summary["move_ms_total"],
            "fill_ms_total": summary["fill_ms_total"], "harvest_ms_total": summary["harvest_ms_total"],
            "idle_wait_ms_total": summary["idle_wait_ms_total"],
            "admission_ms_p50": summary["admission_ms_p50"], "steps": summary["steps"], "wall_s": wall,
            "device_busy_s": summary["steps"] and sum(srv.stats.step_ms) * 1e-3}


def fp64_error(cfg, ps, g):
    """Error of the GPU's first density and force pass against the FP64 oracle on the same input:
    density relative, derivatives relative to the sum of |pair terms| (the FP32 summation bound)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref

    o = oracle_ref.OracleRun(cfg, ps, "f64", scales=True)
    o.resort()
    o.density()
    o.forces()
    k = o.by_id()
    t = (o.species[k] == 0) & (o.mask[k] == 1)

    def worst(gv, ov, sc):
        e = np.abs(np.asarray(gv, np.float64) - ov) / (sc + 1e-300)
        return float(e[..., t].max())

    out = {"rho": worst(g.rho, o.rho[k], o.rho[k]), "dvdt": worst(g.dvdt, o.dvdt[:, k], o.sc_a[k]),
           "dudt": worst(g.dudt, o.dudt[k], o.sc_du[k])}
    if cfg.mhd:
        out["dbdt"] = worst(g.dbdt, o.dbdt[:, k], o.sc_db[k])
        out["divb"] = worst(g.divb, o.divb[k], o.sc_divb[k])
    # FP32 pair coordinates put a few pairs within rounding of r = H on the other side of the cull than
    # FP64 does; the bit-exact count contract is with the oracle of the same precision
    out["neighbour_counts_equal_fp64"] = bool(np.array_equal(g.nnb, o.nnb[k]))
    out["neighbour_count_mismatches_fp64"] = int(np.count_nonzero(g.nnb.astype(np.int64) - o.nnb[k]))
    return out


_T0 = time.perf_counter()


def _progress(phase: str):
    """One line per finished phase on stderr (the JSON result stays the only stdout line)."""
    print(f"[bench {time.perf_counter() - _T0:6.1f} s] {phase} done", file=sys.stderr, flush=True)


def _time_steps(run, budget_s: float):
    steps = 0
    t0 = time.perf_counter()
    while True:
        run.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 20:
            return steps, el


def cpu_baseline(cfg, ps, budget_s: float):
    """The binned CPU restatement: the GPU path's own algorithm (27-supercell stencil over the re-sorted
    state) in C, OpenMP over supercells."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref

    o = oracle_ref.OracleRun(cfg, ps, "mixed")
    o.prime()
    steps, el = _time_steps(o, budget_s)
    return {"value": ps.n * steps / el, "unit": "particle-steps/s", "cores": int(oracle_ref.lib().orc_num_threads()),
            "kind": "port",
            "sample": f"{steps} full steps of the same {ps.n:,}-particle workload (C restatement of the "
                      f"reference CPU arithmetic on the binned 27-supercell stencil, OpenMP over supercells), "
                      f"{el:.1f} s"}


def cpu_reference_structured(cfg, ps, budget_s: float, block: int):
    """The reference's own CPU algorithm restated in C: regions fixed at set-up in 64-slot frames, AABB
    refit, brute-force region-pair search into CSR, every slot of every neighbour frame swept, one
    OpenMP iteration per target frame (NeighbourRegions.hpp:55-135,
    ParticleParticleInteraction.hpp:219-314); FP32 absolute coordinates as the reference's Real."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref

    r = oracle_ref.RegionRun(cfg, ps, "plain", block=block)
    r.prime()
    steps, el = _time_steps(r, budget_s)
    return {"value": ps.n * steps / el, "unit": "particle-steps/s", "cores": int(oracle_ref.lib().orc_num_threads()),
            "kind": "port",
            "sample": f"{steps} full steps of the same {ps.n:,}-particle workload with the reference's region "
                      f"traversal: {r.n_regions:,} regions of {block}^3 supercells in 64-slot frames "
                      f"({r.n:,} slots), brute-force AABB search ({r.csr_entries:,} region pairs), frame sweep, "
                      f"FP32, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128, help="lattice points per axis (128 -> 2,097,152)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp32"])
    ap.add_argument("--tune", action="store_true", help="run the autotuner before timing")
    ap.add_argument("--tuning", default="", help="JSON launch shape for the main workload, e.g. "
                    "'{\"pair_balance\": 2, \"density_chunk\": 512, \"force_chunk\": 256}'")
    ap.add_argument("--cell-factor", type=float, default=0.0,
                    help="supercell edge in interaction radii for the main workload (0: the default, 1; 0.5: the "
                         "125-supercell stencil)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of each CPU figure")
    ap.add_argument("--cpu-region-block", type=int, default=2,
                    help="supercells per region edge of the reference-structured CPU figure (2 was the fastest "
                         "of 1, 2, 3 at 2,097,152 particles: the region search is quadratic in the region count)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-runs", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=50)
    ap.add_argument("--no-tune-ensemble", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-budget-gb", type=float, default=4.0)
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--serve-host-ic", action="store_true", help="serve host-generated particles instead of device ICs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (the step would shard by "
              "supercell bricks with halo exchange; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    m = run_main(args)
    n, K = m["n"], args.steps
    value = n * K / m["wall"]
    f_ms = m["prof"]["stage_ms"]["forces"]
    pairs = m["pairs"]
    force_flop = FORCE_FLOP_PER_PAIR * pairs + FORCE_FLOP_PER_PARTICLE * n
    achieved = force_flop / (f_ms * 1e-3) / 1e12
    out = {
        "metric": "SPH-MHD particle-steps/s at 2M particles; ensemble p50/p99 step latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": m["wall"] * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 positions/reductions" if args.precision == "mixed" else "f32",
        "data": "synthetic (seeded jittered lattice, Gaussian random velocity and B fields)",
        "config": {"workload": f"MHD turbulence {args.n}^3 periodic cube, {args.precision} precision, "
                               "Wendland C2 ~57 neighbours, viscosity + SPMHD + cleaning, CFL dt",
                   "particles": n, "precision": args.precision, "tuning": m["tuning"]},
        "stage_ms": m["prof"]["stage_ms"],
        "step_ms_p50": m["prof"]["step_ms_p50"],
        "step_ms_p99": m["prof"]["step_ms_p99"],
        "neighbours_mean": pairs / n,
        "roofline": {
            "kernel": "k_forces (force pass)",
            "bound": "mfma",
            "compute_path": "FP32 VALU (pair physics is no contraction, MFMA idle); priced against the "
                            "published FP32 peak",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": measured_traffic("k_forces", n, m["tuning"]),
            "flop_per_launch": force_flop,
            "launch_ms": f_ms,
            "algorithmic_bytes_per_launch": FORCE_BYTES_PER_PARTICLE * n,
            "hbm_frac": FORCE_BYTES_PER_PARTICLE * n / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "conserved": m["cons"],
    }
    out["step_roofline"] = step_roofline(n, pairs, m["prof"]["step_ms_mean"])
    full = m["full"]
    out["configured_run"] = {
        "steps": CONFIG_STEPS, "ms_per_step_mean": full["step_ms_mean"], "step_ms_p50": full["step_ms_p50"],
        "step_ms_p99": full["step_ms_p99"], "stage_ms": full["stage_ms"],
        "particle_steps_per_s": n * CONFIG_STEPS / full["wall_s"],
        "neighbours_mean_last_step": full["pairs_last"] / n,
        "note": "the same initial state primed and stepped 100 times: the configuration's whole run"}
    _progress(f"main workload: {value:.4g} particle-steps/s")
    if not args.no_ensemble:
        out["ensemble"] = run_ensemble(args)
        _progress("ensemble")
    if not args.no_serve:
        out["serving"] = run_serving(args)
        _progress("serving")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_reference_structured(m["cfg"], m["ps"], args.cpu_seconds, args.cpu_region_block)
        _progress("CPU reference-structured")
        out["cpu_baseline_binned"] = cpu_baseline(m["cfg"], m["ps"], args.cpu_seconds)
        _progress("CPU binned")
        out["fp64_oracle_error"] = fp64_error(m["cfg"], m["ps"], m["primed"])
        out["vs_cpu"] = value / out["cpu_baseline"]["value"]
        out["vs_cpu_binned"] = value / out["cpu_baseline_binned"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
