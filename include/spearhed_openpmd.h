/*
 * spearhed_openpmd.h -- openPMD particle output of the MI355X SPH-MHD step (HDF5 backend).
 *
 * Replaces spearHED's openPMD plugin (include/spearhed/plugins/openPMD/OpenPMDOutput.hpp:45-166,
 * OpenPMDPlugin.x.cpp notify): one openPMD 1.1.0 file per iteration ("file-based" encoding, the
 * reference's default pattern output_%06T.h5), particles grouped by species, one record per
 * attribute with the reference's record names and SI unit dimensions
 * (include/spearhed/plugins/openPMD/ParticleTraits.hpp): id, position (x, y, z; absolute
 * coordinates), mass, velocity (x, y, z), mass_density, specific_internal_energy, plus
 * magnetic_field (x, y, z) and psi for MHD runs and the standard's positionOffset (constant zero).
 * unitSI is 1 for every component: code units are taken as SI, as the reference does.
 *
 * Species: live particles of SPH_SPECIES_DEFAULT are written as "fluid" (the species name the
 * reference's writer and its Sod validator use), SPH_SPECIES_BOUNDARY as "boundary",
 * SPH_SPECIES_TRACER as "tracer"; empty species are omitted, dead particles (mask 0) are not
 * written. Particles keep the order of the input arrays.
 *
 * This library (libspearhed_openpmd.so) is separate from the step library so that the step does not
 * depend on HDF5. Return codes are those of spearhed_sph.h; sph_openpmd_last_error() gives the
 * message of the last failure in the calling thread.
 */
#ifndef SPEARHED_OPENPMD_H
#define SPEARHED_OPENPMD_H

#include <stdint.h>

#include "spearhed_sph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* File name of `iteration` under `pattern`: %T is replaced by the iteration number, %0<w>T by it
 * zero-padded to w digits (the reference's default is "output_%06T.h5"). Writes at most `cap` bytes
 * including the terminating zero; returns the length, or -1 if the pattern has no %T token or the
 * buffer is too small. */
int32_t sph_openpmd_file_name(const char* pattern, uint64_t iteration, char* out, int32_t cap);

/* Write host columns as one openPMD iteration. Required columns: id, mask, species, pos, mass, vel,
 * density, internal_energy; with `mhd`, bfield and psi as well. `time` and `dt` are the iteration's
 * attributes (simulation time and step). */
int sph_openpmd_write_soa(const sph_host_soa* soa, int32_t mhd, const char* pattern, uint64_t iteration, double time,
                          double dt);

/* Download run `run` of a context (sph_download) and write it as one openPMD iteration, with the
 * run's simulated time and its current dt (the CFL step it takes next) as the iteration's time and
 * dt. Synchronises the context's stream. */
int sph_write_openpmd(sph_ctx* ctx, int32_t run, const char* pattern, uint64_t iteration);

const char* sph_openpmd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SPEARHED_OPENPMD_H */
