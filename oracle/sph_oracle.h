/*
 * sph_oracle.h -- CPU oracle for the SPH(-MHD) particle step. TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of spearHED's CPU-backend arithmetic for the particle step,
 * extended with the physics the reference lacks, and it is the checker that the HIP path is
 * compared against. Only tests, the smoke check and the benchmark's CPU-baseline leg may load it;
 * the product library never links, loads or calls it.
 *
 * Each function follows a reference file:line (cited at the function). The body is written once
 * in sph_oracle_impl.h and compiled twice: REAL=float (suffix _f32; precision modes FP32-plain
 * and mixed) and REAL=double (suffix _f64; the FP64 oracle, absolute FP64 coordinates).
 *
 * Pinning: every function is checked by tests/test_oracle_*.py against the reference's
 * known-answer tests (DensitySummation.cpp, HydroForces.cpp, Kernels.cpp,
 * ParticleParticleInteraction.cpp, and for the region traversal FindNeighbourRegions.cpp and
 * RegionBoundsUpdate.cpp), closed forms (kernel normalisation and derivatives, two-body forces,
 * linear-shear induction, uniform-state invariants), conservation identities, and a brute-force
 * O(N^2) neighbour search. No function is left unpinned.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>
#include "../include/spearhed_sph.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PREC_PLAIN 0  /* FP32 absolute coordinates (SPH_PREC_FP32) */
#define ORC_PREC_MIXED 1  /* FP32 supercell-relative coordinates, FP64 positions (SPH_PREC_MIXED) */
#define ORC_PREC_F64 2    /* FP64 everything (only with the _f64 functions) */

typedef struct orc_grid {
    int32_t n[3];
    int32_t ncell;
    int32_t support;
    double lo[3], L[3], s[3], inv_s[3];
    float lo32[3], L32[3], s32[3], inv_s32[3];
    float H, H2;
    double Hd, H2d;
    int32_t reach;  /* supercells the stencil reaches per side: 1 (edge >= H, 27 cells) or 2 (edge >= H/2,
                       125 cells) */
} orc_grid;

/* Particle state of one run. REAL arrays are void* so one struct serves both precisions:
 * float32 arrays for the _f32 functions, float64 arrays for the _f64 functions. */
typedef struct orc_state {
    int64_t n;
    uint64_t* id;
    uint8_t* mask;
    uint8_t* species;
    double* pos[3];
    void* mass;
    void* vel[3];
    void* rho;
    void* h;
    void* u;
    void* B[3];
    void* psi;
    void* dvdt[3];
    void* dudt;
    void* dbdt[3];
    void* dpsidt;
    void* divb;
    void* vsig;
    uint32_t* nnb;
    /* optional (may be NULL): sum over pairs of |term| for each accumulated field, in FP64 --
       the scale that a summation rounding error is bounded by */
    double* sc_rho;
    double* sc_a;
    double* sc_du;
    double* sc_db;
    double* sc_dpsi;
    double* sc_divb;
    /* run scalars */
    double dt;    /* dt of the next step */
    double time;
    double c_h;   /* cleaning speed of the last density pass */
} orc_state;

int orc_make_grid(const sph_run_config* cfg, orc_grid* g);
/* Supercell key of every particle (same rules as the HIP path; bit-exact contract). */
int orc_bin(const sph_run_config* cfg, int32_t prec, int64_t n, double* const pos[3], uint32_t* key);
/* Sort permutation by (key, id): sorted[i] = old[perm[i]]; cell_start has ncell+1 entries. */
int orc_sort(int64_t n, const uint32_t* key, const uint64_t* id, int32_t ncell, int64_t* perm,
             uint32_t* cell_start);

#define ORC_DECL(S)                                                                                 \
    int orc_permute_##S(orc_state* st, const int64_t* perm);                                        \
    int orc_density_##S(const sph_run_config* cfg, int32_t prec, orc_state* st,                     \
                        const uint32_t* cell_start);                                                \
    int orc_forces_##S(const sph_run_config* cfg, int32_t prec, orc_state* st,                      \
                       const uint32_t* cell_start);                                                 \
    int orc_drift_##S(const sph_run_config* cfg, int32_t prec, orc_state* st, double dt);          \
    int orc_kick_##S(const sph_run_config* cfg, int32_t prec, orc_state* st, double dt,            \
                     sph_conserved* out);                                                           \
    int orc_reduce_##S(const sph_run_config* cfg, int32_t prec, orc_state* st, sph_conserved* out); \
    int orc_resort_##S(const sph_run_config* cfg, int32_t prec, orc_state* st, uint32_t* cell_start); \
    int orc_prime_##S(const sph_run_config* cfg, int32_t prec, orc_state* st, sph_conserved* out); \
    int orc_step_##S(const sph_run_config* cfg, int32_t prec, orc_state* st, int32_t n_steps,      \
                     sph_conserved* out);                                                           \
    int orc_brute_counts_##S(const sph_run_config* cfg, int32_t prec, orc_state* st);              \
    int orc_regions_step_##S(const sph_run_config* cfg, int32_t prec, orc_state* st, int32_t n_regions, \
                             const int64_t* region_start, int32_t n_steps, int64_t* csr_entries,   \
                             sph_conserved* out);                                                   \
    int orc_regions_refit_##S(const sph_run_config* cfg, int32_t prec, const orc_state* st,         \
                              int32_t n_regions, const int64_t* region_start, void* box);         \
    int64_t orc_region_search_##S(int32_t nt, const void* tbox, int32_t ns, const void* sbox,      \
                                  double expand, const double L[3], const int32_t periodic[3],     \
                                  int64_t* off, int32_t* ent, uint8_t* img);                       \
    double orc_kernel_w_##S(int32_t kernel, double r, double h);                                    \
    double orc_kernel_dwdr_##S(int32_t kernel, double r, double h);

ORC_DECL(f32)
ORC_DECL(f64)

/* The reference's region/frame traversal (the CPU baseline's reference-structured algorithm):
 * orc_regions_layout assigns each particle a slot in 64-slot frames of supercell-block regions;
 * orc_regions_step_S steps a state laid out that way (n_steps = 0: the priming evaluation). */
int orc_regions_layout(const sph_run_config* cfg, int32_t prec, int64_t n, double* const pos[3], const uint64_t* id,
                       int32_t block, int64_t* slot, int64_t* region_start, int32_t cap_regions, int32_t* n_regions,
                       int64_t* n_slots);

int orc_num_threads(void);
int orc_thread_num(void); /* index of the calling OpenMP thread, < orc_num_threads() */

#ifdef __cplusplus
}
#endif
#endif
