"""ctypes binding of the C ABI declared in include/spearhed_sph.h.

The structures below mirror the header field by field. ``load()`` opens the in-tree
libspearhed_sph.so and raises if it is missing: there is no CPU fallback for the particle step.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspearhed_sph.so")

SPH_OK = 0
SPH_E_INVALID, SPH_E_NOMEM, SPH_E_DEVICE, SPH_E_STATE, SPH_E_CAPACITY = -1, -2, -3, -4, -5
ERRORS = {
    -1: "SPH_E_INVALID",
    -2: "SPH_E_NOMEM",
    -3: "SPH_E_DEVICE",
    -4: "SPH_E_STATE",
    -5: "SPH_E_CAPACITY",
}

KERNEL_CUBIC, KERNEL_QUINTIC, KERNEL_WENDLAND = 0, 1, 2
PREC_FP32, PREC_MIXED = 0, 1
SPECIES_DEFAULT, SPECIES_BOUNDARY, SPECIES_TRACER = 0, 1, 2


class RunConfig(C.Structure):
    _fields_ = [
        ("box_lo", C.c_double * 3),
        ("box_len", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
        ("kernel", C.c_int32),
        ("h0", C.c_float),
        ("gamma", C.c_float),
        ("alpha_visc", C.c_float),
        ("beta_visc", C.c_float),
        ("mhd", C.c_int32),
        ("cleaning", C.c_int32),
        ("borve", C.c_int32),
        ("clean_sigma", C.c_float),
        ("cfl", C.c_float),
        ("cfl_force", C.c_float),
        ("dt_fixed", C.c_double),
        ("cell_factor", C.c_float),
        ("reserved", C.c_int32),
    ]


def make_config(box_lo=(0.0, 0.0, 0.0), box_len=(1.0, 1.0, 1.0), periodic=(1, 1, 1), kernel=KERNEL_WENDLAND,
                h0=0.05, gamma=5.0 / 3.0, alpha_visc=1.0, beta_visc=2.0, mhd=0, cleaning=0, borve=0,
                clean_sigma=1.0, cfl=0.3, cfl_force=0.25, dt_fixed=0.0, cell_factor=1.0) -> RunConfig:
    c = RunConfig()
    c.box_lo[:] = list(box_lo)
    c.box_len[:] = list(box_len)
    c.periodic[:] = [int(p) for p in periodic]
    c.kernel = kernel
    c.h0 = h0
    c.gamma = gamma
    c.alpha_visc = alpha_visc
    c.beta_visc = beta_visc
    c.mhd = int(mhd)
    c.cleaning = int(cleaning)
    c.borve = int(borve)
    c.clean_sigma = clean_sigma
    c.cfl = cfl
    c.cfl_force = cfl_force
    c.dt_fixed = dt_fixed
    c.cell_factor = cell_factor
    c.reserved = 0
    return c


def config_dict(c: RunConfig) -> dict:
    return {f: (list(getattr(c, f)) if hasattr(getattr(c, f), "__len__") else getattr(c, f)) for f, _ in c._fields_}


_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


SETUP_UNIFORM, SETUP_ROTOR, SETUP_ORSZAG_TANG, SETUP_SHEAR_LAYER = 0, 1, 2, 3


class LatticeBlock(C.Structure):
    _fields_ = [
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("spacing", C.c_double),
        ("density", C.c_double),
        ("pressure", C.c_double),
        ("velocity", C.c_float * 3),
        ("bfield", C.c_float * 3),
        ("species", C.c_int32),
        ("setup", C.c_int32),
        ("params", C.c_double * 6),
    ]


def make_block(lo, hi, spacing, density, pressure, velocity=(0.0, 0.0, 0.0), bfield=(0.0, 0.0, 0.0), species=0,
               setup=SETUP_UNIFORM, params=()) -> LatticeBlock:
    b = LatticeBlock()
    b.lo[:] = [float(v) for v in lo]
    b.hi[:] = [float(v) for v in hi]
    b.spacing, b.density, b.pressure = float(spacing), float(density), float(pressure)
    b.velocity[:] = [float(v) for v in velocity]
    b.bfield[:] = [float(v) for v in bfield]
    b.species, b.setup = int(species), int(setup)
    b.params[:] = [float(v) for v in params] + [0.0] * (6 - len(params))
    return b


class HostSoA(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("id", C.POINTER(C.c_uint64)),
        ("mask", C.POINTER(C.c_uint8)),
        ("species", C.POINTER(C.c_uint8)),
        ("pos", _dp * 3),
        ("mass", _fp),
        ("vel", _fp * 3),
        ("density", _fp),
        ("smoothing_length", _fp),
        ("internal_energy", _fp),
        ("bfield", _fp * 3),
        ("psi", _fp),
        ("dvdt", _fp * 3),
        ("dudt", _fp),
        ("dbdt", _fp * 3),
        ("dpsidt", _fp),
        ("divb", _fp),
        ("vsig", _fp),
        ("n_neighbours", C.POINTER(C.c_uint32)),
    ]


class Conserved(C.Structure):
    _fields_ = [
        ("mass", C.c_double),
        ("momentum", C.c_double * 3),
        ("energy", C.c_double),
        ("dt_next", C.c_double),
        ("time", C.c_double),
        ("c_h", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {"mass": self.mass, "momentum": list(self.momentum), "energy": self.energy,
                "dt_next": self.dt_next, "time": self.time, "c_h": self.c_h}


class GridInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int32 * 3),
        ("n_cells", C.c_int32),
        ("cell", C.c_double * 3),
        ("radius", C.c_float),
        ("particle_offset", C.c_int64),
        ("cell_offset", C.c_int64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("density_chunk", C.c_int32),
        ("force_chunk", C.c_int32),
        ("waves_per_cu", C.c_int32),
        ("max_lanes_per_target", C.c_int32),
        ("pair_balance", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


NUM_STAGES = 5
STAGE_NAMES = ("drift_bin", "resort", "density", "forces", "integrate")


class Profile(C.Structure):
    _fields_ = [
        ("steps", C.c_int32),
        ("stage_ms", C.c_double * NUM_STAGES),
        ("step_ms_mean", C.c_double),
        ("step_ms_p50", C.c_double),
        ("step_ms_p99", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {"steps": self.steps, "stage_ms": dict(zip(STAGE_NAMES, self.stage_ms)),
                "step_ms_mean": self.step_ms_mean, "step_ms_p50": self.step_ms_p50,
                "step_ms_p99": self.step_ms_p99}


NOALLOC_MAX_PARTICLES = 262144
NOALLOC_MAX_CELLS = 262144


class DeviceSoA(C.Structure):
    """sph_device_soa: device pointers (as integers) of one run's columns for sph_density_noalloc."""
    _fields_ = [
        ("n", C.c_int64),
        ("mask", C.c_void_p),
        ("species", C.c_void_p),
        ("pos", C.c_void_p * 3),
        ("pos32", C.c_void_p * 3),
        ("mass", C.c_void_p),
        ("smoothing_length", C.c_void_p),
    ]


# name -> (restype, argtypes); exactly the entry points of include/spearhed_sph.h
_CTX = C.c_void_p
PROTOTYPES = {
    "sph_ctx_create": (C.c_int, [C.POINTER(RunConfig), C.POINTER(C.c_int64), C.c_int32, C.c_int32, C.c_void_p,
                                 C.POINTER(C.c_void_p)]),
    "sph_ctx_destroy": (C.c_int, [_CTX]),
    "sph_pool_create": (C.c_int, [C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                  C.POINTER(C.c_void_p)]),
    "sph_pool_set_runs": (C.c_int, [_CTX, C.POINTER(C.c_int32), C.c_int32, C.POINTER(RunConfig),
                                    C.POINTER(C.c_int64), C.c_int32]),
    "sph_last_error": (C.c_char_p, [_CTX]),
    "sph_sync": (C.c_int, [_CTX]),
    "sph_mark": (C.c_int, [_CTX, C.POINTER(C.c_uint64)]),
    "sph_mark_wait": (C.c_int, [_CTX, C.c_uint64]),
    "sph_get_grid": (C.c_int, [_CTX, C.c_int32, C.POINTER(GridInfo)]),
    "sph_get_run": (C.c_int, [_CTX, C.c_int32, C.POINTER(RunConfig), C.POINTER(C.c_int64)]),
    "sph_run_count": (C.c_int32, [_CTX]),
    "sph_set_tuning": (C.c_int, [_CTX, C.POINTER(Tuning)]),
    "sph_get_tuning": (C.c_int, [_CTX, C.POINTER(Tuning)]),
    "sph_bytes_per_run": (C.c_int64, [C.POINTER(RunConfig), C.c_int64, C.c_int32]),
    "sph_run_cells": (C.c_int64, [C.POINTER(RunConfig)]),
    "sph_pool_bytes": (C.c_int64, [C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_int32]),
    "sph_lattice_sites": (C.c_int64, [C.POINTER(LatticeBlock), C.c_int32]),
    "sph_download_async": (C.c_int, [_CTX, C.c_int32, C.POINTER(HostSoA), C.POINTER(Conserved)]),
    "sph_host_alloc": (C.c_int, [C.c_int64, C.POINTER(C.c_void_p)]),
    "sph_host_free": (C.c_int, [C.c_void_p]),
    "sph_init_run": (C.c_int, [_CTX, C.c_int32, C.POINTER(LatticeBlock), C.c_int32, C.c_uint64, C.c_double]),
    "sph_upload": (C.c_int, [_CTX, C.c_int32, C.POINTER(HostSoA)]),
    "sph_download": (C.c_int, [_CTX, C.c_int32, C.POINTER(HostSoA)]),
    "sph_stage_drift": (C.c_int, [_CTX, C.c_double]),
    "sph_stage_resort": (C.c_int, [_CTX]),
    "sph_stage_neighbour_counts": (C.c_int, [_CTX]),
    "sph_stage_density": (C.c_int, [_CTX]),
    "sph_stage_forces": (C.c_int, [_CTX]),
    "sph_stage_integrate": (C.c_int, [_CTX, C.c_double]),
    "sph_reduce_min_dt": (C.c_int, [_CTX, C.POINTER(C.c_double)]),
    "sph_reduce_conserved": (C.c_int, [_CTX, C.POINTER(Conserved)]),
    "sph_prime": (C.c_int, [_CTX]),
    "sph_step": (C.c_int, [_CTX, C.c_int32]),
    "sph_graph_capture": (C.c_int, [_CTX]),
    "sph_graph_launch": (C.c_int, [_CTX, C.c_int32]),
    "sph_get_sort": (C.c_int, [_CTX, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "sph_profile_begin": (C.c_int, [_CTX, C.c_int32]),
    "sph_profile_end": (C.c_int, [_CTX, C.POINTER(Profile), C.POINTER(C.c_double)]),
    "sph_autotune": (C.c_int, [_CTX, C.c_int32, C.POINTER(Tuning)]),
    "sph_probe_copy": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "sph_density_noalloc": (C.c_int, [C.POINTER(RunConfig), C.c_int32, C.POINTER(DeviceSoA), C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_char_p, C.c_int32]),
}

_lib = None


class SphError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Open the gfx950 library. Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SphError(f"{path} is missing: build it with `python -m spearhed_amd.build` "
                       "(the SPH step runs only on the HIP library)")
    lib = C.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


OPENPMD_LIB_PATH = os.path.join(HERE, "libspearhed_openpmd.so")
OPENPMD_PROTOTYPES = {
    "sph_openpmd_file_name": (C.c_int32, [C.c_char_p, C.c_uint64, C.c_char_p, C.c_int32]),
    "sph_openpmd_write_soa": (C.c_int, [C.POINTER(HostSoA), C.c_int32, C.c_char_p, C.c_uint64, C.c_double,
                                        C.c_double]),
    "sph_write_openpmd": (C.c_int, [_CTX, C.c_int32, C.c_char_p, C.c_uint64]),
    "sph_openpmd_last_error": (C.c_char_p, []),
}
_openpmd = None


def load_openpmd(path: str = OPENPMD_LIB_PATH):
    """Open the openPMD writer (include/spearhed_openpmd.h), after the step library it links to."""
    global _openpmd
    if _openpmd is not None:
        return _openpmd
    load()
    if not os.path.exists(path):
        raise SphError(f"{path} is missing: build it with `python -m spearhed_amd.build` (needs HDF5)")
    lib = C.CDLL(path)
    for name, (res, args) in OPENPMD_PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _openpmd = lib
    return lib


def check_openpmd(rc: int, what: str = "") -> None:
    if rc != SPH_OK:
        msg = load_openpmd().sph_openpmd_last_error().decode(errors="replace")
        raise SphError(f"{what}: {ERRORS.get(rc, rc)} {msg}")


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc != SPH_OK:
        msg = ""
        if ctx:
            msg = load().sph_last_error(ctx).decode(errors="replace")
        raise SphError(f"{what}: {ERRORS.get(rc, rc)} {msg}")
