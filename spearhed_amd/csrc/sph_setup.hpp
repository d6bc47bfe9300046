// sph_setup.hpp -- device initial conditions (internal to the library; see sph_setup.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/spearhed_sph.h"
#include "sph_device.hpp"

namespace sph {

constexpr int kMaxSetupBlocks = 8;

__host__ __device__ uint32_t sc_truncate_near_integer(double v);
__host__ __device__ void sc_counts(const sph_lattice_block& b, uint32_t n[3]);
__host__ __device__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1);

// Particles placed by the blocks, or -1 if a block is invalid (host).
int64_t lattice_sites(const sph_lattice_block* blocks, int n_blocks);
hipError_t launch_init_lattice(const sph_lattice_block* blocks, int n_blocks, int64_t np, int64_t p0, uint64_t seed,
                               double jitter, float h0, float gamma, int mixed, int has_mhd, const StateBuf& S,
                               hipStream_t stream);

}  // namespace sph
