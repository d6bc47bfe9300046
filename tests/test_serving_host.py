"""Host-side logic of the continuous-batching server (CPU): FIFO admission under the HBM budget,
retirement after each request's steps, and the run moves handed to the pool. The device pool is
replaced by a recorder; the GPU tests cover the pool itself."""
from __future__ import annotations

import pytest

import ctypes as C

from spearhed_amd import capi, ic
from spearhed_amd.serving import Request, Server, poisson_arrivals, pool_capacity


def _budget_for(particles, max_runs=512, cpp=0.12):
    """The reservation of a pool of `particles` capacity (the server's default cell ratio)."""
    cells = max(64, int(-(-cpp * particles // 1)))
    return int(capi.load().sph_pool_bytes(int(particles), cells, max_runs, capi.PREC_MIXED, 1))


class RecorderPool:
    """Stands in for stepper.Pool: records the calls the server makes and keeps the run table."""

    def __init__(self, max_particles, max_cells, max_runs, precision, mhd):
        self.lib = capi.load()
        self.sizes, self.calls, self.steps = [], [], 0
        self.max_particles, self.max_cells, self.max_runs = max_particles, max_cells, max_runs

    def cells_needed(self, cfg):
        import ctypes as C
        return int(self.lib.sph_run_cells(C.byref(cfg)))

    def set_runs(self, keep, add):
        assert len(set(keep)) == len(keep) and all(0 <= k < len(self.sizes) for k in keep)
        self.sizes = [self.sizes[k] for k in keep] + [n for _, n in add]
        assert sum(self.sizes) <= self.max_particles and len(self.sizes) <= self.max_runs
        self.calls.append(("set_runs", list(keep), len(add)))
        return list(range(len(keep), len(self.sizes)))

    def upload(self, run, ps):
        assert ps.n == self.sizes[run]
        self.calls.append(("upload", run))

    def prime(self):
        self.calls.append(("prime",))

    def step(self, n):
        self.steps += n

    def sync(self):
        pass

    def mark(self):
        self.calls.append(("mark",))
        return len(self.calls)

    def wait_mark(self, ticket):
        assert 0 < ticket <= len(self.calls)

    def download(self, run):
        return ("state", run, self.sizes[run])

    def conserved(self):
        return [{"run": r} for r in range(len(self.sizes))]

    def close(self):
        pass


def _requests(k, steps, n_target=2000):
    out = []
    for r in range(k):
        cfg, ps = ic.kelvin_helmholtz(n_target, 300 + r)
        out.append(Request(r, cfg, ps, steps[r % len(steps)], arrival=float(r)))
    return out


@pytest.mark.parametrize("budget_gb", [0.01, 0.5, 4.0, 64.0])
def test_pool_reservation_fits_the_budget(budget_gb):
    budget = int(budget_gb * 2 ** 30)
    srv = Server(budget, max_runs=256, pool_factory=RecorderPool)
    reserved = capi.load().sph_pool_bytes(srv.max_particles, srv.max_cells, srv.max_runs, capi.PREC_MIXED, 1)
    assert srv.pool_bytes == reserved
    assert 0 < reserved <= budget
    # the largest such pool: one more particle of capacity does not fit
    p, c = pool_capacity(budget, 256, 0.12)
    assert (p, c) == (srv.max_particles, srv.max_cells)
    assert _budget_for(p + 1, 256) > budget
    assert reserved > 0.99 * budget - 64 * 2 ** 10


def test_budget_limits_the_resident_set():
    reqs = _requests(6, [3, 5])
    srv = Server(_budget_for(int(2.5 * reqs[0].ps.n)), pool_factory=RecorderPool)
    res_sizes = []
    orig = srv._retire_and_admit

    def spy(until):
        orig(until)
        res_sizes.append(len(srv.resident))
        assert sum(r.req.n for r in srv.resident) <= srv.max_particles
        assert srv.pool_bytes <= srv.budget

    srv._retire_and_admit = spy
    srv.serve(reqs, realtime=False)
    assert max(res_sizes) == 2  # two runs fit the budget, a third does not
    assert sorted(srv.results) == list(range(6))
    for q in reqs:
        assert srv.results[q.rid].steps == q.steps


def test_fifo_and_retire_order():
    reqs = _requests(4, [2, 4])
    srv = Server(1 << 34, pool_factory=RecorderPool)
    srv.serve(reqs, realtime=False)
    admits = [c for c in srv.pool.calls if c[0] == "set_runs"]
    assert admits[0] == ("set_runs", [], 4)           # all fit: admitted together
    assert admits[1] == ("set_runs", [1, 3], 0)       # after their priming step + 2 steps runs 0, 2 retire
    assert srv.pool.steps == 5                        # the longest request: priming step + 4 steps
    assert ("prime",) not in srv.pool.calls           # admitted runs are primed by their first step
    assert srv.results[0].ps[1] == 0 and srv.results[2].ps[1] == 2  # downloaded before the move


def test_request_larger_than_budget_is_rejected():
    reqs = _requests(1, [1])
    srv = Server(_budget_for(reqs[0].ps.n - 300), pool_factory=RecorderPool)  # 300: beyond 256-byte padding
    with pytest.raises(capi.SphError):
        srv.submit(reqs[0])


def test_poisson_arrivals_rate():
    t = poisson_arrivals(20000, 50.0, seed=1)
    assert (t[1:] >= t[:-1]).all()
    assert abs(len(t) / t[-1] - 50.0) < 2.0


class DepthPool(RecorderPool):
    """A recorder that completes marks in order and records how many steps were in flight (launched,
    not yet waited for) each time a step is launched."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.outstanding, self.depths = [], []

    def step(self, n):
        self.depths.append(len(self.outstanding))
        super().step(n)

    def mark(self):
        t = super().mark()
        self.outstanding.append(t)
        return t

    def wait_mark(self, ticket):
        super().wait_mark(ticket)
        self.outstanding = [t for t in self.outstanding if t > ticket]

    def sync(self):
        self.outstanding = []


def test_steps_in_flight_stay_within_the_depth():
    from spearhed_amd.serving import STEPS_IN_FLIGHT, RESPONSE_RINGS
    assert STEPS_IN_FLIGHT < RESPONSE_RINGS  # a ring's copy-outs are harvested before it is refilled
    reqs = _requests(6, [3, 5, 2])
    srv = Server(1 << 34, pool_factory=DepthPool)
    srv.serve(reqs, realtime=False)
    d = srv.pool.depths
    assert max(d) == STEPS_IN_FLIGHT  # the pipeline fills, and never beyond its depth
    assert srv.pool.outstanding == []
    assert sorted(srv.results) == [r.rid for r in reqs]
    for r in reqs:
        assert srv.results[r.rid].steps == r.steps  # after its priming step
