"""sph_density_noalloc (the density pass for callers that cannot allocate device memory) against the
context path it is a variant of: sph_upload + sph_stage_resort + sph_stage_density + sph_download of
the same particles, ids 0..n-1 in input order. Density and neighbour counts must be bit-identical.

Two calls in a row on one GPU share the library's scratch array; the second, smaller call must not
see the first one's leftovers. The first call is at the maximum size (64^3 = 262,144 particles).
The caller's device buffers are allocated here with the HIP runtime the library itself links (the
test is the caller; the entry point allocates nothing).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from parity import gpu_stages
from spearhed_amd import capi, ic

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    capi.load()  # the library first: the runtime it links is the one already in the process
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
        getattr(hip, f).restype = C.c_int
    return hip


class _Dev:
    """Device copies of host arrays, freed on close."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(arr.nbytes, 1)) == 0
        self.ptrs.append(p.value)
        assert self.hip.hipMemcpy(p.value, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p.value

    def get(self, ptr, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(cfg, ps, precision, pass_flags=True):
    hip = _hip()
    dev = _Dev(hip)
    try:
        soa = capi.DeviceSoA()
        soa.n = ps.n
        if pass_flags:
            soa.mask = dev.put(ps.mask.astype(np.uint8))
            soa.species = dev.put(ps.species.astype(np.uint8))
        for a in range(3):
            if precision == "mixed":
                soa.pos[a] = dev.put(ps.pos[a].astype(np.float64))
            else:
                soa.pos32[a] = dev.put(ps.pos[a].astype(np.float32))
        soa.mass = dev.put(ps.mass.astype(np.float32))
        soa.smoothing_length = dev.put(ps.h.astype(np.float32))
        rho = dev.put(np.full(ps.n, -1.0, dtype=np.float32))
        nnb = dev.put(np.full(ps.n, 7, dtype=np.uint32))
        err = C.create_string_buffer(256)
        prec = capi.PREC_MIXED if precision == "mixed" else capi.PREC_FP32
        rc = capi.load().sph_density_noalloc(C.byref(cfg), prec, C.byref(soa), rho, nnb, None, err, 256)
        assert rc == capi.SPH_OK, err.value.decode()
        assert hip.hipDeviceSynchronize() == 0
        return dev.get(rho, ps.n, np.float32), dev.get(nnb, ps.n, np.uint32)
    finally:
        dev.close()


def _context(cfg, ps, precision):
    q = ps.copy()
    q.id = np.arange(q.n, dtype=np.uint64)  # the no-alloc path orders ties by input index
    q.rho[:] = 0.0  # what the no-alloc path reports for particles that are not targets
    g = gpu_stages(cfg, q, precision, stages=("resort", "density"))
    return g.rho, g.nnb


def _case(name):
    if name == "turbulence_max":
        return ic.mhd_turbulence(n=64, seed=11)  # 262,144 particles: the maximum
    if name == "turbulence_small":
        return ic.mhd_turbulence(n=20, seed=12)
    if name == "sod_walls":
        return ic.sod_tube(seed=13, nx_left=32, cross=6)  # walls: Boundary species outside the box
    raise KeyError(name)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_noalloc_two_calls_match_context(precision):
    for name in ("turbulence_max", "turbulence_small"):
        cfg, ps = _case(name)
        print(f"noalloc {name} {precision} n={ps.n}", flush=True)
        assert ps.n <= capi.NOALLOC_MAX_PARTICLES
        rho, nnb = _noalloc(cfg, ps, precision, pass_flags=(name == "turbulence_max"))
        rho_c, nnb_c = _context(cfg, ps, precision)
        assert np.array_equal(nnb, nnb_c), f"{name}: neighbour counts differ"
        assert np.array_equal(rho.view(np.uint32), rho_c.view(np.uint32)), f"{name}: density differs"
        assert nnb.min() > 20 and np.all(rho > 0)


def test_noalloc_half_h_supercells_match_context():
    """Supercells of edge H/2 (cell_factor 0.5): the no-alloc pass takes the 125-supercell stencil as
    the context does, with the same bits."""
    cfg, ps = _case("turbulence_small")
    cfg.cell_factor = 0.5
    print(f"noalloc half-H n={ps.n}", flush=True)
    rho, nnb = _noalloc(cfg, ps, "mixed")
    rho_c, nnb_c = _context(cfg, ps, "mixed")
    assert np.array_equal(nnb, nnb_c) and np.array_equal(rho.view(np.uint32), rho_c.view(np.uint32))
    assert nnb.min() > 20


def test_noalloc_walls_report_targets_only():
    cfg, ps = _case("sod_walls")
    print(f"noalloc sod_walls n={ps.n}", flush=True)
    rho, nnb = _noalloc(cfg, ps, "mixed")
    rho_c, nnb_c = _context(cfg, ps, "mixed")
    tgt = (ps.species == 0) & (ps.mask == 1)
    assert (~tgt).sum() > 0
    assert np.array_equal(nnb, nnb_c) and np.array_equal(rho.view(np.uint32), rho_c.view(np.uint32))
    assert np.all(rho[~tgt] == 0) and np.all(nnb[~tgt] == 0)


def test_noalloc_rejects_oversize_before_launching():
    cfg = capi.make_config(h0=0.05)
    soa = capi.DeviceSoA()
    soa.n = capi.NOALLOC_MAX_PARTICLES + 1
    err = C.create_string_buffer(256)
    rc = capi.load().sph_density_noalloc(C.byref(cfg), capi.PREC_MIXED, C.byref(soa), 1, 1, None, err, 256)
    assert rc == capi.SPH_E_CAPACITY and b"maximum" in err.value
