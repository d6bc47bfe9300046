"""Continuous-batching server of independent SPH runs on one GPU.

Requests (a run configuration, its initial particles and a number of steps) arrive over time and
wait in a FIFO queue. The server keeps one device pool (sph_pool_create) whose whole reservation,
sph_pool_bytes(particles, supercells, runs), fits the HBM budget: the largest particle capacity P
with supercell capacity ceil(cells_per_particle * P) that does. Before every batched step it retires
runs that have finished their steps (their final state is downloaded) and admits waiting requests in
arrival order while the resident runs fit the pool's particle, supercell and run capacities.

All resident runs then advance one step together (one launch per stage for the whole batch). Every
run keeps its own dt and time, and its summation orders depend only on its own data, so a run's
result is bit-identical whether it is served alone or in a mix.

The step latency recorded per batched step is its device time (HIP events on the pool's stream); the
host-side interval between consecutive step completions is kept beside it (one step is kept in
flight, so the host's admission work overlaps the device's step). p50 / p99 come from those samples.
"""
from __future__ import annotations

import collections
import ctypes as C
import heapq
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import capi
from .capi import RunConfig
from .particles import ParticleSet
from .stepper import Pool


RESPONSE_RINGS = 3  # pinned response rings, each sized for the pool's whole capacity
# Batched steps kept in flight while the host retires and admits runs. A ring's copy-outs are
# harvested once the step they follow completes, STEPS_IN_FLIGHT rounds after they were enqueued, and
# a ring is refilled RESPONSE_RINGS retiring rounds after its last use: the depth must stay below
# the ring count.
STEPS_IN_FLIGHT = 2


def pool_capacity(budget: int, max_runs: int, cells_per_particle: float, precision: str = "mixed",
                  mhd: bool = True):
    """(particles, supercells) of the largest pool whose reservation sph_pool_bytes fits `budget`,
    with supercells = ceil(cells_per_particle * particles) (at least 64). The reservation grows with
    the particle capacity, so the largest fitting one is found by bisection."""
    lib = capi.load()
    prec = capi.PREC_MIXED if precision == "mixed" else capi.PREC_FP32

    def cells(p):
        return max(64, int(math.ceil(cells_per_particle * p)))

    def fits(p):
        b = int(lib.sph_pool_bytes(int(p), cells(p), int(max_runs), prec, int(mhd)))
        return 0 < b <= budget

    lo = 1024
    if not fits(lo):
        raise capi.SphError(f"an HBM budget of {budget} bytes cannot hold a pool of {lo} particles")
    hi = max(lo, budget // 64)  # no particle costs less than 64 bytes
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if fits(mid):
            lo = mid
        else:
            hi = mid - 1
    return lo, cells(lo)


def run_bytes(cfg: RunConfig, n: int, precision: str = "mixed") -> int:
    """Device bytes a run of n particles costs in the admission budget (sph_bytes_per_run)."""
    lib = capi.load()
    b = int(lib.sph_bytes_per_run(C.byref(cfg), int(n), capi.PREC_MIXED if precision == "mixed" else capi.PREC_FP32))
    if b < 0:
        raise capi.SphError("invalid run configuration")
    return b


@dataclass
class Request:
    """A run to serve: its configuration and either its particles (uploaded from the host) or lattice
    blocks (capi.LatticeBlock, filled on the device by sph_init_run with `seed`)."""
    rid: int
    cfg: RunConfig
    ps: Optional[ParticleSet]
    steps: int
    arrival: float = 0.0  # seconds after the server's start
    blocks: Optional[list] = None
    seed: int = 0
    n: int = field(init=False, default=0)

    def __post_init__(self):
        if self.ps is not None:
            self.n = self.ps.n
        else:
            arr = (capi.LatticeBlock * len(self.blocks))(*self.blocks)
            self.n = int(capi.load().sph_lattice_sites(arr, len(self.blocks)))
            if self.n < 0:
                raise capi.SphError(f"request {self.rid}: invalid lattice blocks")


@dataclass
class Result:
    rid: int
    ps: Optional[ParticleSet]  # full final state (full downloads) or None
    conserved: dict
    steps: int
    arrival: float
    admitted: float
    finished: float
    fields: dict = field(default_factory=dict)  # selected response fields (copy-out path)


@dataclass
class _Resident:
    req: Request
    done: int = 0
    admitted: float = 0.0


@dataclass
class ServeStats:
    step_ms: List[float] = field(default_factory=list)
    batch_runs: List[int] = field(default_factory=list)
    batch_particles: List[int] = field(default_factory=list)
    admit_ms: List[float] = field(default_factory=list)
    retire_ms: List[float] = field(default_factory=list)   # reductions + downloads of finished runs
    move_ms: List[float] = field(default_factory=list)     # sph_pool_set_runs
    fill_ms: List[float] = field(default_factory=list)     # uploads / device fills of admitted runs
    step_wall_ms: List[float] = field(default_factory=list)  # host-side interval between step completions
    harvest_ms: List[float] = field(default_factory=list)    # responses turned into results
    idle_ms: List[float] = field(default_factory=list)       # no run resident: waiting for an arrival

    def summary(self) -> dict:
        s = np.array(self.step_ms) if self.step_ms else np.zeros(1)
        ps = np.array(self.batch_particles, dtype=np.float64) if self.batch_particles else np.zeros(1)
        busy = s.sum() * 1e-3
        return {
            "steps": len(self.step_ms),
            "step_ms_p50": float(np.percentile(s, 50)),
            "step_ms_p99": float(np.percentile(s, 99)),
            "step_wall_ms_p50": float(np.percentile(self.step_wall_ms, 50)) if self.step_wall_ms else None,
            "step_wall_ms_p99": float(np.percentile(self.step_wall_ms, 99)) if self.step_wall_ms else None,
            "mean_batch_runs": float(np.mean(self.batch_runs)) if self.batch_runs else 0.0,
            "particle_steps_per_s": float(ps.sum() / busy) if busy > 0 else 0.0,
            "admission_ms_total": float(np.sum(self.admit_ms)) if self.admit_ms else 0.0,
            "retire_ms_total": float(np.sum(self.retire_ms)) if self.retire_ms else 0.0,
            "move_ms_total": float(np.sum(self.move_ms)) if self.move_ms else 0.0,
            "fill_ms_total": float(np.sum(self.fill_ms)) if self.fill_ms else 0.0,
            "harvest_ms_total": float(np.sum(self.harvest_ms)) if self.harvest_ms else 0.0,
            "admission_ms_p50": float(np.percentile(self.admit_ms, 50)) if self.admit_ms else 0.0,
            "idle_wait_ms_total": float(np.sum(self.idle_ms)) if self.idle_ms else 0.0,
        }


class Server:
    """Continuous batching over a fixed device pool.

    hbm_budget: device bytes the server may use: the pool's whole reservation fits it.
    cells_per_particle: supercells reserved per particle of capacity (a uniform run with ~57
        neighbours uses about 1/14 per particle; a slab one supercell thick about 1/20; a Brio-Wu tube
        with its eightfold density jump about 1/9).
    """

    def __init__(self, hbm_budget: int, precision: str = "mixed", max_runs: int = 512, mhd: bool = True,
                 cells_per_particle: float = 0.12, pool_factory=Pool, response_fields=None):
        """response_fields: None returns each run's full final state (synchronous download); a tuple of
        names (stepper.RESPONSE_FIELDS) returns only those, copied out asynchronously through pinned
        memory while the pool moves on."""
        self.lib = capi.load()
        self.precision = precision
        self.prec_code = capi.PREC_MIXED if precision == "mixed" else capi.PREC_FP32
        self.budget = int(hbm_budget)
        self.max_runs = int(max_runs)
        self.max_particles, self.max_cells = pool_capacity(self.budget, self.max_runs, cells_per_particle,
                                                           precision, mhd)
        self.pool = pool_factory(self.max_particles, self.max_cells, self.max_runs, precision, mhd)
        self.pool_bytes = int(self.lib.sph_pool_bytes(self.max_particles, self.max_cells, self.max_runs,
                                                      self.prec_code, int(mhd)))
        self.resident: List[_Resident] = []
        self.queue: List[tuple] = []  # (arrival, rid, Request)
        self.stats = ServeStats()
        self.results: dict = {}
        self._t0 = None
        self.response_fields = tuple(response_fields) if response_fields else None
        # Responses are copied out into pinned rings used in turn by retiring rounds: a round's
        # copies complete with the step behind which they were enqueued and are then copied on into
        # the results by worker threads, while the next retiring rounds fill the other rings.
        self._rings = []
        self._turn = 0
        self._pending = []  # (ring, resident, arrays, conserved, finished time): copies enqueued
        self._ready = []    # the same, after a synchronisation: safe to read
        if self.response_fields:
            from .stepper import RESPONSE_FIELDS, PinnedRing
            per = sum(np.dtype(RESPONSE_FIELDS[f][1]).itemsize * RESPONSE_FIELDS[f][2]
                      for f in self.response_fields)
            ring_bytes = self.max_particles * per + 512 * (len(self.response_fields) + 1) * 256
            self._rings = [PinnedRing(ring_bytes) for _ in range(RESPONSE_RINGS)]
        # responses leave the pinned rings on worker threads (numpy copies release the GIL), so the
        # loop goes on stepping; a ring is reused only after its copies have finished
        self._copiers = ThreadPoolExecutor(max_workers=8) if self.response_fields else None
        self._copies = [[] for _ in self._rings]  # outstanding copy jobs per ring

    # -- bookkeeping -------------------------------------------------------------------------------
    def run_bytes(self, req: Request) -> int:
        return run_bytes(req.cfg, req.n, self.precision)

    def resident_bytes(self) -> int:
        return sum(self.run_bytes(r.req) for r in self.resident)

    def submit(self, req: Request):
        if req.n > self.max_particles or self.pool.cells_needed(req.cfg) > self.max_cells:
            raise capi.SphError(f"request {req.rid} does not fit the server's HBM budget")
        heapq.heappush(self.queue, (req.arrival, req.rid, req))

    def now(self) -> float:
        return time.perf_counter() - self._t0

    # -- one scheduling round ----------------------------------------------------------------------
    def _retire_and_admit(self, admit_until: float):
        """Retire finished runs, then admit queued requests that arrived by admit_until."""
        now = self.now()
        keep, finished = [], []
        for i, r in enumerate(self.resident):
            (finished if r.done >= r.req.steps else keep).append(i)
        t_r = time.perf_counter()
        if self.response_fields:
            from .stepper import download_async
            ring = self._turn
            if finished:
                self._turn = (self._turn + 1) % len(self._rings)
                self._drain(ring)  # the ring's last copies out must be done before it is refilled
            for i in finished:
                arrays, cons = download_async(self.pool, i, self.response_fields, self._rings[ring])
                self._pending.append((ring, self.resident[i], arrays, cons, now))
        else:
            cons = self.pool.conserved() if finished else []
            for i in finished:
                r = self.resident[i]
                self.results[r.req.rid] = Result(r.req.rid, self.pool.download(i), cons[i], r.done, r.req.arrival,
                                                 r.admitted, now)
        self.stats.retire_ms.append((time.perf_counter() - t_r) * 1e3)
        used_p = sum(self.resident[i].req.n for i in keep)
        used_c = sum(self.pool.cells_needed(self.resident[i].req.cfg) for i in keep)
        add: List[Request] = []
        while self.queue and self.queue[0][0] <= admit_until and len(keep) + len(add) < self.max_runs:
            req = self.queue[0][2]
            n, c = req.n, self.pool.cells_needed(req.cfg)
            if used_p + n > self.max_particles or used_c + c > self.max_cells:
                break  # FIFO: the head waits until enough retires
            heapq.heappop(self.queue)
            add.append(req)
            used_p, used_c = used_p + n, used_c + c
        if not finished and not add:
            return
        t_m = time.perf_counter()
        new_idx = self.pool.set_runs(keep, [(q.cfg, q.n) for q in add])  # enqueued behind the steps in flight
        self.stats.move_ms.append((time.perf_counter() - t_m) * 1e3)
        t_f = time.perf_counter()
        # an admitted run's first batched step primes it (density, forces, first dt; no motion),
        # so it needs one step more than it asked for and no separate priming pass
        self.resident = [self.resident[i] for i in keep] + [_Resident(q, -1, now) for q in add]
        for j, q in zip(new_idx, add):
            if q.ps is not None:
                self.pool.upload(j, q.ps)
            else:
                self.pool.init_run(j, q.blocks, seed=q.seed)
        self.stats.fill_ms.append((time.perf_counter() - t_f) * 1e3)

    def _synced(self):
        """The pool's stream has been synchronised: every copy-out enqueued so far is complete."""
        self._ready += self._pending
        self._pending = []

    def _harvest(self):
        """Turn completed asynchronous copy-outs into results: each response's arrays are copied out of
        the pinned ring on the worker threads (a Result's fields are complete once the serve call
        returns, or once the ring has been drained)."""
        if not self._ready:
            return
        t0 = time.perf_counter()
        for ring, r, arrays, cons, fin in self._ready:
            fields = {}
            for k, v in arrays.items():
                dst = np.empty_like(v)
                fields[k] = dst
                self._copies[ring].append(self._copiers.submit(np.copyto, dst, v))
            self.results[r.req.rid] = Result(r.req.rid, None, cons.as_dict(), r.done, r.req.arrival, r.admitted,
                                             fin, fields=fields)
        self._ready = []
        self.stats.harvest_ms.append((time.perf_counter() - t0) * 1e3)

    def _drain(self, ring: int):
        """Wait for the copies out of a ring, then free it for new copy-outs."""
        t0 = time.perf_counter()
        for f in self._copies[ring]:
            f.result()
        self._copies[ring] = []
        self._rings[ring].reset()
        self.stats.harvest_ms.append((time.perf_counter() - t0) * 1e3)

    def serve(self, requests: Optional[List[Request]] = None, max_steps: int = 1 << 30,
              realtime: bool = True) -> dict:
        """Run until every request has finished. realtime=False admits a request as soon as the
        queue reaches it (arrival times only order the queue).

        STEPS_IN_FLIGHT steps are kept in flight: each round retires and admits runs (copy-outs,
        run-set change and device fills enqueued behind the steps in flight), enqueues the next
        batched step, and only then waits for the oldest step beyond that depth, so that the host's
        admission work overlaps the device's steps even when a round takes longer than one step.
        step_ms is each step's device time (HIP events around it); step_wall_ms the host-side
        interval between consecutive step completions."""
        for q in requests or []:
            self.submit(q)
        # per-step device time from events recorded on the pool's stream (host work between steps,
        # such as copying finished responses, does not inflate it)
        profiled = hasattr(self.pool, "profile_begin")
        if profiled:
            self.pool.profile_begin(sum(q[2].steps + 1 for q in self.queue) + 8)
        self._t0 = time.perf_counter()
        n_steps = 0
        flight = collections.deque()  # (mark of a step, copy-outs enqueued before it), oldest first
        last_done = None
        while (self.queue or self.resident or flight) and n_steps < max_steps:
            t_adm = time.perf_counter()
            self._retire_and_admit(self.now() if realtime else float("inf"))
            self.stats.admit_ms.append((time.perf_counter() - t_adm) * 1e3)
            if self.resident:
                self.pool.step(1)
                flight.append((self.pool.mark(), self._pending))
                self._pending = []
                self.stats.batch_runs.append(len(self.resident))
                self.stats.batch_particles.append(sum(r.req.n for r in self.resident))
                for r in self.resident:
                    r.done += 1
                n_steps += 1
            # wait for the oldest steps beyond the depth (and what preceded them); with no run
            # resident, for all of them
            while flight and (len(flight) > STEPS_IN_FLIGHT or not self.resident):
                mark, pending = flight.popleft()
                self.pool.wait_mark(mark)
                now = time.perf_counter()
                if last_done is not None:
                    self.stats.step_wall_ms.append((now - last_done) * 1e3)
                last_done = now
                self._ready += pending
                self._harvest()  # copied while the steps still in flight run
            if not self.resident:
                self.pool.sync()
                self._synced()
                self._harvest()
                last_done = None
                if self.queue:  # idle until the next arrival
                    t_idle = time.perf_counter()
                    time.sleep(max(0.0, min(self.queue[0][0] - self.now(), 0.01)))
                    self.stats.idle_ms.append((time.perf_counter() - t_idle) * 1e3)
        while flight:  # stopped by max_steps with steps in flight: their copy-outs stay pending
            self._pending = flight.pop()[1] + self._pending
        self._retire_and_admit(float("-inf"))  # retire the last runs, admit nothing
        self.pool.sync()
        self._synced()
        self._harvest()
        for ring in range(len(self._rings)):
            self._drain(ring)
        if profiled:
            prof = self.pool.profile_end()
            self.stats.step_ms = list(prof["step_ms"])
        return self.stats.summary()

    def close(self):
        self.pool.close()
        if self._copiers is not None:
            self._copiers.shutdown(wait=True)
        for ring in self._rings:
            ring.close()


def poisson_arrivals(n: int, rate_per_s: float, seed: int = 0) -> np.ndarray:
    """Arrival times of a Poisson process (exponential gaps) with the given mean rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))
