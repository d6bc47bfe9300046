// sph_device.hpp -- device-side data layout and SPH arithmetic of the gfx950 particle step.
//
// Layout in HBM (one context = one or more runs, concatenated):
//   * particle state, structure of arrays, sorted by (supercell key, particle id) after every
//     resort; two buffer sets (A/B) so the resort gathers without an in-place permutation;
//   * packed interaction records (float4 planes) written once per pass and staged into LDS:
//       P0 = (x, y, z, m)   pair coordinate (FP32 absolute or supercell-relative) and mass;
//                           non-sources carry a sentinel position so the cull rejects them
//       P1 = (vx, vy, vz, c)         c = fast magnetosonic bound sqrt((gamma P + B^2)/rho)
//       P2 = (h, 1/h, gfac, rho)     gfac = kernel-gradient prefactor
//       P3 = (pr, ph, bdr, psr)      pr = (P + B^2/2)/rho^2, ph = P/rho^2, bdr = 1/rho^2, psr = psi/rho^2
//       P4 = (Bx, By, Bz, psi)       (MHD only)
//   * supercell table: per-cell particle start (exclusive scan of the per-cell histogram).
//
// The arithmetic mirrors the reference's functors: CubicSplineKernel.hpp:47-137,
// QuinticSplineKernel.hpp:46-160 (shape functions and the gradWScalar guard), DensitySummation.hpp:50-94,
// HydroForces.hpp:112-175, EquationOfState.hpp:32-47; Wendland C2 per Dehnen & Aly (2012).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

// Per-run constants, uploaded once per composition change.
struct RunDev {
    int64_t p0, np;          // first particle, particle count
    int32_t c0, ncell;       // first supercell, supercell count
    int32_t n[3];            // supercells per axis
    int32_t periodic[3];
    double lo[3], L[3], s[3], inv_s[3];
    float lo32[3], L32[3], s32[3], inv_s32[3];
    float H, H2;             // interaction radius and its square (FP32, as used by the cull)
    int32_t kernel, mhd, cleaning, support;
    float gamma, alpha, beta, borve, clean_sigma, cfl, cfl_force, sigma;
    double dt_fixed;
    int32_t kb_first, kb_count;  // particle-kernel blocks of this run
    float f0;                    // kernel shape at q = 0
    int32_t reach;               // supercells the stencil reaches per side: 1 (edge >= H) or 2 (edge >= H/2)
};

// Device arrays of one particle buffer set.
struct StateBuf {
    uint64_t* id;
    uint8_t* mask;
    uint8_t* species;
    double* pos64[3];  // mixed precision
    float* pos32[3];   // plain FP32
    float* mass;
    float* h;
    float* u;
    float* rho;
    float* psi;
    float* vel[3];
    float* B[3];
};

struct WorkBuf {
    float* dvdt[3];
    float* dudt;
    float* dbdt[3];
    float* dpsidt;
    float* divb;
    float* vsig;
    uint32_t* nnb;
    uint32_t* key;
    uint32_t* rank;
    uint32_t* tmp;
    uint64_t* tmpid;  // particle id staged next to tmp by the scatter (gather ranks by it)
    uint32_t* nbmask; // accepted-candidate bits per supercell-wave lane: word w of lane l of cell c
                      // at nbmask[(c * kMaskWords + w) * 64 + l]
    uint32_t* count;
    uint32_t* cs;     // ncell_tot + 1 entries
    uint32_t* bsum;
    float4* P0;
    float4* P1;
    float4* P2;
    float4* P3;
    float4* P4;
    double* dt;       // per run: dt of the current step
    double* time;     // per run
    uint32_t* chbits; // per run: max fast speed (float bits, atomicMax)
    double* partial;  // per kick block: 6 doubles (dt, M, Px, Py, Pz, E)
    double* cons;     // per run: 8 doubles (mass, px, py, pz, energy, dt_next, time, c_h)
    int32_t* blk_run; // per particle-kernel block: run id
    int32_t* prime_flag; // per run: 1 = the next non-advancing reduction sets the run's first dt
    int64_t* move_tab;   // per kept run when a pool's run set changes: destination start, source start, count
    int32_t* run_src;    // per run of a new run set: the old run it continues, or -1 for a new run
    double* run_tmp;     // per old run: its dt, time and priming flag, saved while the run set changes
    int32_t* perm;       // optional (may be null): the resort writes each sorted slot's source index here
};

constexpr float kSentinel = 8.507059e37f;  // FLT_MAX / 4: sentinel^2 overflows to +inf, so r2 < H2 fails
// Mask words kept per lane: the density pass stores which of a lane's first 32*kMaskWords
// candidates it accepted, and the force pass replays them instead of culling again.
constexpr int kMaskWords = 8;

__device__ __forceinline__ bool is_target(uint8_t sp, uint8_t live) { return live && sp == 0; }
__device__ __forceinline__ bool is_source(uint8_t sp, uint8_t live) { return live && sp <= 1; }
__device__ __forceinline__ bool is_movable(uint8_t sp, uint8_t live) { return live && (sp == 0 || sp == 2); }

// Kernel shape f(q) with W = sigma/h^3 f(q). The smoothing-kernel functions are written with
// explicit fused multiply-adds and every other contraction off, so that the instantiation that reads
// the kernel per run (batches mixing kernels) rounds exactly as the one that fixes it at compile time:
// a run's results do not depend on the company it is batched with.
__device__ __forceinline__ float kshape(int kern, float q)
{
#pragma clang fp contract(off)
    if (kern == 2) {  // Wendland C2: (1 - q/2)^4 (1 + 2q)
        // 1 - q/2 clamped at 0: exactly 0 for q >= 2 (the support test), the same bits below it
        const float t = fmaxf(fmaf(-0.5f, q, 1.0f), 0.0f);
        const float t2 = t * t;
        return (t2 * t2) * fmaf(2.0f, q, 1.0f);
    }
    if (kern == 0) {  // cubic M4: 1 - 3/2 q^2 + 3/4 q^3 inside, (2 - q)^3 / 4 outside
        const float inner = fmaf(q * q, fmaf(0.75f, q, -1.5f), 1.0f);
        const float t = 2.0f - q;
        const float outer = 0.25f * ((t * t) * t);
        return q < 1.0f ? inner : (q < 2.0f ? outer : 0.0f);
    }
    // quintic M5: (3 - q)^5 - 6 (2 - q)^5 + 15 (1 - q)^5, terms dropped beyond their support
    const float a = 3.0f - q, b = 2.0f - q, c = 1.0f - q;
    const float a2 = a * a, b2 = b * b, c2 = c * c;
    const float a5 = (a2 * a2) * a, b5 = (b2 * b2) * b, c5 = (c2 * c2) * c;
    if (q < 1.0f) return fmaf(15.0f, c5, fmaf(-6.0f, b5, a5));
    if (q < 2.0f) return fmaf(-6.0f, b5, a5);
    if (q < 3.0f) return a5;
    return 0.0f;
}

// Gradient scalar s with grad_i W_ij = s r_ij (gfac includes sigma and the h powers).
__device__ __forceinline__ float kgrad(int kern, float rr, float invR, float q, float gfac)
{
#pragma clang fp contract(off)
    float s;
    if (kern == 2) {
        const float t = fmaf(-0.5f, q, 1.0f);
        s = gfac * ((t * t) * t);
        s = q < 2.0f ? s : 0.0f;
    } else if (kern == 0) {
        const float gi = q * fmaf(2.25f, q, -3.0f);
        const float t = 2.0f - q;
        const float go = -0.75f * (t * t);
        const float g = q < 1.0f ? gi : (q < 2.0f ? go : 0.0f);
        s = (gfac * g) * invR;
    } else {
        const float a = 3.0f - q, b = 2.0f - q, c = 1.0f - q;
        const float a2 = a * a, b2 = b * b, c2 = c * c;
        const float a4 = a2 * a2, b4 = b2 * b2, c4 = c2 * c2;
        const float g = q < 1.0f ? fmaf(-75.0f, c4, fmaf(30.0f, b4, -5.0f * a4))
                                 : (q < 2.0f ? fmaf(30.0f, b4, -5.0f * a4) : (q < 3.0f ? -5.0f * a4 : 0.0f));
        s = (gfac * g) * invR;
    }
    return rr > 0.0f ? s : 0.0f;
}

__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// r from r^2 in one instruction (v_sqrt_f32, about 1 ulp): sqrt(0) = 0, so coincident particles
// need no guard, where r2 * rsq(r2) does (0 * inf)
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Squared distance from the own pair coordinate to a staged (pre-shifted) neighbour, as
// fma(dz, dz, fma(dy, dy, dx*dx)) with every other contraction off: the cull decision and the
// neighbour counts are part of the bit-exact contract, and correctly rounded fused operations give
// the same bits on any IEEE machine.
__device__ __forceinline__ float geom_r2(float xi, float yi, float zi, float4 p, float& dx, float& dy, float& dz)
{
#pragma clang fp contract(off)
    dx = xi - p.x;
    dy = yi - p.y;
    dz = zi - p.z;
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));  // three roundings, reproduced exactly by the oracle
}

}  // namespace sph
