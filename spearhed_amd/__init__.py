"""spearhed_amd: MI355X (gfx950) SPH-MHD particle step behind a C ABI.

The compute path is libspearhed_sph.so (hand-written HIP kernels); this package holds its ctypes
binding (capi), the host mirror of the reference's stage interface (stepper), a particle container
(particles), initial conditions (ic), the batched ensemble and serving drivers (ensemble, serving)
and the build script (build).
"""
from .capi import (KERNEL_CUBIC, KERNEL_QUINTIC, KERNEL_WENDLAND, PREC_FP32, PREC_MIXED, SPECIES_BOUNDARY,
                   SPECIES_DEFAULT, SPECIES_TRACER, RunConfig, SphError, make_config)
from .particles import ParticleSet

__all__ = ["RunConfig", "make_config", "ParticleSet", "SphError", "KERNEL_CUBIC", "KERNEL_QUINTIC",
           "KERNEL_WENDLAND", "PREC_FP32", "PREC_MIXED", "SPECIES_DEFAULT", "SPECIES_BOUNDARY", "SPECIES_TRACER"]
