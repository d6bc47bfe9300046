// sph_setup.hip -- initial conditions generated on the device (sph_init_run).
//
// A run is filled from up to kMaxSetupBlocks lattice blocks: a simple cubic lattice per block with
// the reference's count and placement rules
// (packages/pmacc/include/spmacc/particles/initialization/SC.hpp:50-59 truncateNearInteger,
// :118-146 computeSCCellCounts, :178-189 cell-centre placement, x fastest), a per-axis uniform
// jitter from Philox4x32-10 keyed by (seed, particle id) (Salmon, Moraes, Dror & Shaw, SC'11), and
// the block's setup for the fields. Positions are computed in FP64 with contraction off, so a CPU
// restatement reproduces them bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>
#include <cmath>

#include "../../include/spearhed_sph.h"
#include "sph_setup.hpp"

namespace sph {

__host__ __device__ uint32_t sc_truncate_near_integer(double v)
{
    const uint32_t t = (uint32_t)v;
    const double scale = v > 1.0 ? v : 1.0;
    const double tol = 4.0 * DBL_EPSILON * scale;
    const double dist = (double)(t + 1u) - v;
    return (dist >= 0.0 && dist <= tol) ? t + 1u : t;
}

__host__ __device__ void sc_counts(const sph_lattice_block& b, uint32_t n[3])
{
    const double ex = b.hi[0] - b.lo[0];
    n[0] = sc_truncate_near_integer(ex / b.spacing);
    if (n[0] == 0u) n[0] = 1u;
    for (int a = 1; a < 3; ++a) {
        const uint32_t v = (uint32_t)((b.hi[a] - b.lo[a]) / b.spacing + 0.5);
        n[a] = v == 0u ? 1u : v;
    }
}

// Philox4x32-10: 10 rounds of (hi, lo) = M * x multiplications with a Weyl-sequence key schedule.
__host__ __device__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

namespace {

struct SetupArgs {
    sph_lattice_block blk[kMaxSetupBlocks];
    int64_t first[kMaxSetupBlocks + 1];  // first particle of each block (run-local)
    uint32_t n[kMaxSetupBlocks][3];
    int nblocks;
    int64_t np, p0;
    uint32_t seed_lo, seed_hi;
    double jitter;
    float h0, gamma;
    int mixed, has_mhd;
};

__global__ __launch_bounds__(256) void k_init_lattice(SetupArgs A, StateBuf S)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= A.np) return;
    int b = 0;
    while (b + 1 < A.nblocks && g >= A.first[b + 1]) ++b;
    const sph_lattice_block& B = A.blk[b];
    const int64_t li = g - A.first[b];
    const uint32_t nx = A.n[b][0], ny = A.n[b][1];
    const int64_t idx[3] = {li % nx, (li / nx) % ny, li / ((int64_t)nx * ny)};
    uint32_t c[4] = {(uint32_t)g, (uint32_t)((uint64_t)g >> 32), 0u, 0u};
    philox4x32_10(c, A.seed_lo, A.seed_hi);
    double x[3], d[3];
    for (int a = 0; a < 3; ++a) {
#pragma clang fp contract(off)
        d[a] = (B.hi[a] - B.lo[a]) / (double)A.n[b][a];
        double v = B.lo[a] + ((double)idx[a] + 0.5) * d[a];
        const double u = ((double)c[a] + 0.5) * 2.3283064365386963e-10;  // (r + 1/2) / 2^32 in (0, 1)
        v = v + A.jitter * d[a] * (2.0 * u - 1.0);
        x[a] = v;
    }
    double rho = B.density;
    double vel[3] = {B.velocity[0], B.velocity[1], B.velocity[2]};
    double bf[3] = {B.bfield[0], B.bfield[1], B.bfield[2]};
    const double twopi = 6.283185307179586;
    if (B.setup == SPH_SETUP_ROTOR) {
        const double rx = x[0] - B.params[0], ry = x[1] - B.params[1];
        const double r = sqrt(rx * rx + ry * ry);
        if (r < B.params[2]) {
            const double om = B.params[3] / B.params[2];
            rho = B.params[4];
            vel[0] = -om * ry;
            vel[1] = om * rx;
        }
    } else if (B.setup == SPH_SETUP_ORSZAG_TANG) {
        vel[0] = -B.params[0] * sin(twopi * x[1]);
        vel[1] = B.params[0] * sin(twopi * x[0]);
        bf[0] = -B.params[1] * sin(twopi * x[1]);
        bf[1] = B.params[1] * sin(2.0 * twopi * x[0]);
    } else if (B.setup == SPH_SETUP_SHEAR_LAYER) {
        const bool band = fabs(x[1] - 0.5) < 0.25;
        rho = band ? B.density * B.params[1] : B.density;
        vel[0] = band ? 0.5 * B.params[0] : -0.5 * B.params[0];
        vel[1] = B.params[2] * sin(2.0 * twopi * x[0] + B.params[3]);
    }
    const double vol = d[0] * d[1] * d[2];
    const int64_t i = A.p0 + g;
    S.id[i] = (uint64_t)g;
    S.mask[i] = 1;
    S.species[i] = (uint8_t)B.species;
    for (int a = 0; a < 3; ++a) {
        if (A.mixed) S.pos64[a][i] = x[a];
        else S.pos32[a][i] = (float)x[a];
        S.vel[a][i] = (float)vel[a];
        if (A.has_mhd) S.B[a][i] = (float)bf[a];
    }
    S.mass[i] = (float)(rho * vol);
    S.h[i] = A.h0;
    S.u[i] = (float)(B.pressure / (((double)A.gamma - 1.0) * rho));
    S.rho[i] = (float)rho;
    if (A.has_mhd) S.psi[i] = 0.0f;
}

}  // namespace

int64_t lattice_sites(const sph_lattice_block* blocks, int n_blocks)
{
    if (!blocks || n_blocks < 1 || n_blocks > kMaxSetupBlocks) return -1;
    int64_t total = 0;
    for (int b = 0; b < n_blocks; ++b) {
        const sph_lattice_block& B = blocks[b];
        if (!(B.spacing > 0.0) || !(B.density > 0.0) || !(B.pressure >= 0.0)) return -1;
        for (int a = 0; a < 3; ++a)
            if (!(B.hi[a] > B.lo[a])) return -1;
        if (B.setup < SPH_SETUP_UNIFORM || B.setup > SPH_SETUP_SHEAR_LAYER) return -1;
        if (B.species < 0 || B.species > 2) return -1;
        uint32_t n[3];
        sc_counts(B, n);
        total += (int64_t)n[0] * n[1] * n[2];
    }
    return total;
}

hipError_t launch_init_lattice(const sph_lattice_block* blocks, int n_blocks, int64_t np, int64_t p0, uint64_t seed,
                               double jitter, float h0, float gamma, int mixed, int has_mhd, const StateBuf& S,
                               hipStream_t stream)
{
    SetupArgs A = {};
    int64_t acc = 0;
    for (int b = 0; b < n_blocks; ++b) {
        A.blk[b] = blocks[b];
        sc_counts(blocks[b], A.n[b]);
        A.first[b] = acc;
        acc += (int64_t)A.n[b][0] * A.n[b][1] * A.n[b][2];
    }
    A.first[n_blocks] = acc;
    A.nblocks = n_blocks;
    A.np = np;
    A.p0 = p0;
    A.seed_lo = (uint32_t)seed;
    A.seed_hi = (uint32_t)(seed >> 32);
    A.jitter = jitter;
    A.h0 = h0;
    A.gamma = gamma;
    A.mixed = mixed;
    A.has_mhd = has_mhd;
    if (np <= 0) return hipSuccess;
    k_init_lattice<<<(unsigned)((np + 255) / 256), 256, 0, stream>>>(A, S);
    return hipGetLastError();
}

}  // namespace sph
