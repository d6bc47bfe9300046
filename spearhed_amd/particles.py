"""Host-side particle container with the reference's attribute set.

Columns follow the fluid record of include/spearhed/param/speciesDefinition.param:100-110
(particleId, multiMask, position, mass, vel, density, smoothingLength, internalEnergy, dvdt, dudt),
extended with the MHD fields (B, psi and their derivatives, div B), the signal speed and the
neighbour count. Vector fields are stored as (3, n) arrays so each component is contiguous (SoA).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import HostSoA

SCALAR_F32 = ("mass", "rho", "h", "u", "psi", "dudt", "dpsidt", "divb", "vsig")
VECTOR_F32 = ("vel", "B", "dvdt", "dbdt")


class ParticleSet:
    """Structure of arrays for one run."""

    def __init__(self, n: int):
        self.n = int(n)
        self.id = np.arange(n, dtype=np.uint64)
        self.mask = np.ones(n, dtype=np.uint8)
        self.species = np.zeros(n, dtype=np.uint8)
        self.pos = np.zeros((3, n), dtype=np.float64)
        for f in SCALAR_F32:
            setattr(self, f, np.zeros(n, dtype=np.float32))
        for f in VECTOR_F32:
            setattr(self, f, np.zeros((3, n), dtype=np.float32))
        self.nnb = np.zeros(n, dtype=np.uint32)

    # -- utilities -------------------------------------------------------------------------------
    def copy(self) -> "ParticleSet":
        p = ParticleSet.__new__(ParticleSet)
        p.n = self.n
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(p, k, v.copy())
        return p

    def take(self, order: np.ndarray) -> "ParticleSet":
        p = ParticleSet.__new__(ParticleSet)
        p.n = self.n
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(p, k, v[..., order].copy())
        return p

    def sorted_by_id(self) -> "ParticleSet":
        return self.take(np.argsort(self.id, kind="stable"))

    @staticmethod
    def concat(parts) -> "ParticleSet":
        p = ParticleSet(0)
        p.n = sum(q.n for q in parts)
        for k, v in parts[0].__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(p, k, np.concatenate([getattr(q, k) for q in parts], axis=-1))
        return p

    # -- C ABI view ------------------------------------------------------------------------------
    def soa(self) -> HostSoA:
        """A HostSoA pointing into this object's arrays (keep the object alive while used)."""
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray) and not v.flags.c_contiguous:
                setattr(self, k, np.ascontiguousarray(v))
        s = HostSoA()
        s.n = self.n

        def fp(a):
            return a.ctypes.data_as(C.POINTER(C.c_float))

        s.id = self.id.ctypes.data_as(C.POINTER(C.c_uint64))
        s.mask = self.mask.ctypes.data_as(C.POINTER(C.c_uint8))
        s.species = self.species.ctypes.data_as(C.POINTER(C.c_uint8))
        for a in range(3):
            s.pos[a] = self.pos[a].ctypes.data_as(C.POINTER(C.c_double))
            s.vel[a] = fp(self.vel[a])
            s.bfield[a] = fp(self.B[a])
            s.dvdt[a] = fp(self.dvdt[a])
            s.dbdt[a] = fp(self.dbdt[a])
        s.mass = fp(self.mass)
        s.density = fp(self.rho)
        s.smoothing_length = fp(self.h)
        s.internal_energy = fp(self.u)
        s.psi = fp(self.psi)
        s.dudt = fp(self.dudt)
        s.dpsidt = fp(self.dpsidt)
        s.divb = fp(self.divb)
        s.vsig = fp(self.vsig)
        s.n_neighbours = self.nnb.ctypes.data_as(C.POINTER(C.c_uint32))
        return s
