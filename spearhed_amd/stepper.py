"""Python host mirror of the reference's stage interface, over the C ABI.

The reference composes its particle step from stage functors (include/spearhed/control/Simulation.x.cpp:196-310):
PushVelocity -> UpdateVolumes -> calculateNeighbours -> UpdateDensity -> UpdateHydroForces ->
EulerIntegrate, followed by one syncWait. The classes below keep those names and that order; each
one is a thin call into libspearhed_sph.so (no arithmetic happens in Python).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import capi
from .capi import Conserved, GridInfo, RunConfig, Tuning, check
from .particles import ParticleSet


class Context:
    """A device-resident set of runs (1 = single run, >1 = ragged batch)."""

    def __init__(self, configs: Sequence[RunConfig], sizes: Sequence[int], precision: str = "mixed"):
        self.lib = capi.load()
        self.nruns = len(configs)
        self.precision = capi.PREC_MIXED if precision == "mixed" else capi.PREC_FP32
        self._cfg = (RunConfig * self.nruns)(*configs)
        self._cap = (C.c_int64 * self.nruns)(*[int(s) for s in sizes])
        self.sizes = [int(s) for s in sizes]
        h = C.c_void_p()
        rc = self.lib.sph_ctx_create(self._cfg, self._cap, self.nruns, self.precision, None, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.sph_last_error(h).decode() if h else ""
            self.close()
            raise capi.SphError(f"sph_ctx_create: {capi.ERRORS.get(rc, rc)} {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.sph_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- data ------------------------------------------------------------------------------------
    def upload(self, run: int, ps: ParticleSet):
        s = ps.soa()
        check(self.lib.sph_upload(self.h, run, C.byref(s)), self.h, "sph_upload")

    def init_run(self, run: int, blocks, seed: int = 0, jitter: float = 0.05):
        """Fill a run on the device from lattice blocks (capi.LatticeBlock); see sph_init_run."""
        arr = (capi.LatticeBlock * len(blocks))(*blocks)
        check(self.lib.sph_init_run(self.h, run, arr, len(blocks), int(seed) & 0xFFFFFFFFFFFFFFFF, float(jitter)),
              self.h, "sph_init_run")

    def download(self, run: int) -> ParticleSet:
        ps = ParticleSet(self.sizes[run])
        s = ps.soa()
        check(self.lib.sph_download(self.h, run, C.byref(s)), self.h, "sph_download")
        return ps

    def grid(self, run: int) -> GridInfo:
        g = GridInfo()
        check(self.lib.sph_get_grid(self.h, run, C.byref(g)), self.h, "sph_get_grid")
        return g

    def sort_info(self, run: int):
        n = self.sizes[run]
        g = self.grid(run)
        key = np.zeros(n, dtype=np.uint32)
        cs = np.zeros(g.n_cells + 1, dtype=np.uint32)
        check(self.lib.sph_get_sort(self.h, run, key.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    cs.ctypes.data_as(C.POINTER(C.c_uint32))), self.h, "sph_get_sort")
        return key, cs

    # -- stages ----------------------------------------------------------------------------------
    def drift(self, dt: float = 0.0):
        check(self.lib.sph_stage_drift(self.h, dt), self.h, "sph_stage_drift")

    def resort(self):
        check(self.lib.sph_stage_resort(self.h), self.h, "sph_stage_resort")

    def neighbour_counts(self):
        check(self.lib.sph_stage_neighbour_counts(self.h), self.h, "sph_stage_neighbour_counts")

    def density(self):
        check(self.lib.sph_stage_density(self.h), self.h, "sph_stage_density")

    def forces(self):
        check(self.lib.sph_stage_forces(self.h), self.h, "sph_stage_forces")

    def integrate(self, dt: float = 0.0):
        check(self.lib.sph_stage_integrate(self.h, dt), self.h, "sph_stage_integrate")

    def prime(self):
        check(self.lib.sph_prime(self.h), self.h, "sph_prime")

    def step(self, n: int = 1):
        check(self.lib.sph_step(self.h, n), self.h, "sph_step")

    def graph_capture(self):
        check(self.lib.sph_graph_capture(self.h), self.h, "sph_graph_capture")

    def graph_launch(self, n: int):
        check(self.lib.sph_graph_launch(self.h, n), self.h, "sph_graph_launch")

    def sync(self):
        check(self.lib.sph_sync(self.h), self.h, "sph_sync")

    def min_dt(self) -> np.ndarray:
        out = (C.c_double * self.nruns)()
        check(self.lib.sph_reduce_min_dt(self.h, out), self.h, "sph_reduce_min_dt")
        return np.array(out[:])

    def write_openpmd(self, run: int, pattern: str, iteration: int):
        """Write run `run` as one openPMD iteration (HDF5, one file per iteration; pattern with %T,
        e.g. "output_%06T.h5"). Returns the file name."""
        lib = capi.load_openpmd()
        capi.check_openpmd(lib.sph_write_openpmd(self.h, run, pattern.encode(), int(iteration)), "sph_write_openpmd")
        buf = C.create_string_buffer(4096)
        lib.sph_openpmd_file_name(pattern.encode(), int(iteration), buf, 4096)
        return buf.value.decode()

    def conserved(self) -> list:
        out = (Conserved * self.nruns)()
        check(self.lib.sph_reduce_conserved(self.h, out), self.h, "sph_reduce_conserved")
        return [o.as_dict() for o in out]

    def profile_begin(self, max_steps: int):
        check(self.lib.sph_profile_begin(self.h, max_steps), self.h, "sph_profile_begin")

    def mark(self) -> int:
        """Ticket of the current point of the context's stream (sph_mark)."""
        t = C.c_uint64()
        check(self.lib.sph_mark(self.h, C.byref(t)), self.h, "sph_mark")
        return t.value

    def wait_mark(self, ticket: int):
        """Block until the stream has passed the point of `ticket` (sph_mark_wait)."""
        check(self.lib.sph_mark_wait(self.h, int(ticket)), self.h, "sph_mark_wait")

    def profile_end(self):
        p = capi.Profile()
        steps = (C.c_double * 100000)()
        check(self.lib.sph_profile_end(self.h, C.byref(p), steps), self.h, "sph_profile_end")
        d = p.as_dict()
        d["step_ms"] = list(steps[: p.steps])
        return d

    def neighbour_total(self, run: int) -> int:
        """Sum of per-particle neighbour counts (accepted pairs) of a run, from the last density pass."""
        from .capi import HostSoA
        nnb = np.zeros(self.sizes[run], dtype=np.uint32)
        s = HostSoA()
        s.n = self.sizes[run]
        s.n_neighbours = nnb.ctypes.data_as(C.POINTER(C.c_uint32))
        check(self.lib.sph_download(self.h, run, C.byref(s)), self.h, "sph_download")
        return int(nnb.sum(dtype=np.uint64))

    @property
    def tuning(self) -> dict:
        t = Tuning()
        check(self.lib.sph_get_tuning(self.h, C.byref(t)), self.h, "sph_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning(**{**self.tuning, **kw})
        check(self.lib.sph_set_tuning(self.h, C.byref(t)), self.h, "sph_set_tuning")

    def autotune(self, n_steps: int = 2) -> dict:
        t = Tuning()
        check(self.lib.sph_autotune(self.h, n_steps, C.byref(t)), self.h, "sph_autotune")
        return t.as_dict()


# host columns of sph_host_soa by response field name: (attribute, dtype, components)
RESPONSE_FIELDS = {
    "id": ("id", np.uint64, 1), "mask": ("mask", np.uint8, 1), "species": ("species", np.uint8, 1),
    "pos": ("pos", np.float64, 3), "vel": ("vel", np.float32, 3), "mass": ("mass", np.float32, 1),
    "rho": ("density", np.float32, 1), "h": ("smoothing_length", np.float32, 1),
    "u": ("internal_energy", np.float32, 1), "B": ("bfield", np.float32, 3), "psi": ("psi", np.float32, 1),
    "dvdt": ("dvdt", np.float32, 3), "dudt": ("dudt", np.float32, 1), "dbdt": ("dbdt", np.float32, 3),
    "dpsidt": ("dpsidt", np.float32, 1), "divb": ("divb", np.float32, 1), "vsig": ("vsig", np.float32, 1),
    "nnb": ("n_neighbours", np.uint32, 1),
}


class PinnedRing:
    """A block of pinned host memory handed out front to back; reset() once every copy into it has
    completed. Buffers for sph_download_async."""

    def __init__(self, nbytes: int):
        self.lib = capi.load()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        if self.lib.sph_host_alloc(self.nbytes, C.byref(p)) != 0:
            raise capi.SphError(f"sph_host_alloc({self.nbytes}) failed")
        self.base = p.value
        self.off = 0

    def take(self, dtype, shape) -> np.ndarray:
        dt = np.dtype(dtype)
        size = int(np.prod(shape)) * dt.itemsize
        start = (self.off + 255) & ~255
        if start + size > self.nbytes:
            raise capi.SphError("pinned ring exhausted")
        self.off = start + size
        buf = (C.c_char * size).from_address(self.base + start)
        return np.frombuffer(buf, dtype=dt).reshape(shape)

    def reset(self):
        self.off = 0

    def close(self):
        if getattr(self, "base", None):
            self.lib.sph_host_free(C.c_void_p(self.base))
            self.base = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def download_async(ctx: "Context", run: int, fields, ring: PinnedRing):
    """Enqueue the copy-out of `fields` (names of RESPONSE_FIELDS) and the run's conserved record into
    `ring`; returns (arrays, conserved): valid once the context's stream has been synchronised."""
    from .capi import HostSoA
    n = ctx.sizes[run]
    s = HostSoA()
    s.n = n
    out = {}
    for f in fields:
        attr, dt, comp = RESPONSE_FIELDS[f]
        arr = ring.take(dt, (comp, n) if comp > 1 else (n,))
        out[f] = arr
        ptype = C.POINTER(np.ctypeslib.as_ctypes_type(dt))
        if comp > 1:
            for a in range(comp):
                getattr(s, attr)[a] = arr[a].ctypes.data_as(ptype)
        else:
            setattr(s, attr, arr.ctypes.data_as(ptype))
    cons_buf = ring.take(np.float64, (8,))
    cons = Conserved.from_address(cons_buf.ctypes.data)
    check(ctx.lib.sph_download_async(ctx.h, run, C.byref(s), C.byref(cons)), ctx.h, "sph_download_async")
    return out, cons


class Pool(Context):
    """A context of fixed capacity whose run set changes between steps (continuous batching).

    Capacities are set once (particles, supercells, runs); set_runs keeps a subset of the resident
    runs, whose state and dt/time move with them, and appends new runs that the caller uploads and
    primes. No device memory is allocated after construction.
    """

    def __init__(self, max_particles: int, max_cells: int, max_runs: int, precision: str = "mixed",
                 mhd: bool = True):
        self.lib = capi.load()
        self.precision = capi.PREC_MIXED if precision == "mixed" else capi.PREC_FP32
        self.max_particles, self.max_cells, self.max_runs = int(max_particles), int(max_cells), int(max_runs)
        self.nruns = 0
        self.sizes: list = []
        self.configs: list = []
        h = C.c_void_p()
        rc = self.lib.sph_pool_create(self.max_particles, self.max_cells, self.max_runs, self.precision,
                                      int(bool(mhd)), None, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.sph_last_error(h).decode() if h else ""
            self.close()
            raise capi.SphError(f"sph_pool_create: {capi.ERRORS.get(rc, rc)} {msg}")

    def set_runs(self, keep: Sequence[int], add: Sequence[tuple]):
        """keep: old run indices, in their new order; add: (RunConfig, n_particles) pairs appended after
        them. Returns the new index of every added run."""
        keep = [int(k) for k in keep]
        k_arr = (C.c_int32 * max(1, len(keep)))(*keep)
        a_cfg = (RunConfig * max(1, len(add)))(*[c for c, _ in add])
        a_n = (C.c_int64 * max(1, len(add)))(*[int(n) for _, n in add])
        check(self.lib.sph_pool_set_runs(self.h, k_arr, len(keep), a_cfg, a_n, len(add)), self.h,
              "sph_pool_set_runs")
        self.configs = [self.configs[k] for k in keep] + [c for c, _ in add]
        self.sizes = [self.sizes[k] for k in keep] + [int(n) for _, n in add]
        self.nruns = len(self.sizes)
        return list(range(len(keep), self.nruns))

    def cells_needed(self, cfg: RunConfig) -> int:
        """Supercells a run of this configuration occupies in the pool (host-side grid arithmetic)."""
        return int(self.lib.sph_run_cells(C.byref(cfg)))


# -- reference-named stage functors ---------------------------------------------------------------
class PushVelocity:
    """r += v dt for Movable species (include/spearhed/particles/pusher/PushVelocity.hpp:35-42)."""

    def __call__(self, ctx: Context, dt: float = 0.0):
        ctx.drift(dt)


class UpdateVolumes:
    """Supercell re-sort; replaces the region AABB refit
    (packages/pmacc/include/spmacc/particles/regions/RegionBoundsUpdate.hpp:175-214)."""

    def __call__(self, ctx: Context):
        ctx.resort()


class UpdateDensity:
    """Self term + pair sum (include/spearhed/particles/density/DensitySummation.hpp:104-138)."""

    def __call__(self, ctx: Context):
        ctx.density()


class UpdateHydroForces:
    """Zero derivatives + pair forces (include/spearhed/sph/HydroForces.hpp:183-214), with viscosity,
    SPMHD and div-B cleaning when the run enables them."""

    def __call__(self, ctx: Context):
        ctx.forces()


class EulerIntegrate:
    """v += dvdt dt, u += dudt dt (+ B, psi) and the FP64 CFL / conservation reductions
    (include/spearhed/particles/pusher/EulerIntegrate.hpp:43-52)."""

    def __call__(self, ctx: Context, dt: float = 0.0):
        ctx.integrate(dt)


class Simulation:
    """Simulation::runOneStep (include/spearhed/control/Simulation.x.cpp:196-310) over a Context.

    openpmd_period > 0 writes every run as an openPMD iteration after each step whose number is a
    multiple of the period, as the reference's OpenPMDPlugin (openPMD.period / openPMD.file options,
    include/spearhed/plugins/openPMD/OpenPMDPlugin.x.cpp) does; with several runs in the context, run
    r goes to the pattern with "_run<r>" before its extension."""

    def __init__(self, ctx: Context, openpmd_period: int = 0, openpmd_file: str = "output_%06T.h5"):
        self.ctx = ctx
        self.push = PushVelocity()
        self.update_volumes = UpdateVolumes()
        self.update_density = UpdateDensity()
        self.update_hydro = UpdateHydroForces()
        self.euler = EulerIntegrate()
        self.openpmd_period = int(openpmd_period)
        self.openpmd_file = openpmd_file
        self.written = []  # files written by the output hook

    def _pattern(self, run: int) -> str:
        if self.ctx.nruns == 1:
            return self.openpmd_file
        stem, dot, ext = self.openpmd_file.rpartition(".")
        return f"{stem}_run{run}.{ext}" if dot else f"{self.openpmd_file}_run{run}"

    def run_one_step(self, dt: float = 0.0, step: int = None):
        self.push(self.ctx, dt)
        self.update_volumes(self.ctx)
        self.update_density(self.ctx)
        self.update_hydro(self.ctx)
        self.euler(self.ctx, dt)
        self.ctx.sync()
        if self.openpmd_period > 0 and step is not None and step % self.openpmd_period == 0:
            for r in range(self.ctx.nruns):
                self.written.append(self.ctx.write_openpmd(r, self._pattern(r), step))
