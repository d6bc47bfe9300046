/*
 * spearhed_sph.h -- C ABI of the MI355X (gfx950) SPH-MHD particle step.
 *
 * This is the drop-in boundary for spearHED's particle step. The reference exposes its stages as
 * header-only C++ functor -> sender templates; this header exposes the same stages as plain C entry
 * points over a device-resident context, so that a C++ (or C / ctypes) host can drive them.
 *
 * Stage <-> reference mapping (paths relative to the spearHED tree):
 *   sph_stage_drift            PushVelocity<S> via launchForEach
 *                              (include/spearhed/particles/pusher/PushVelocity.hpp:35-42,
 *                               packages/pmacc/include/spmacc/particles/algorithms/LaunchForEach.hpp:154-159)
 *   sph_stage_resort           UpdateVolumes<PR> + FrameIndexBuffer::refreshIfStaleSender
 *                              (packages/pmacc/include/spmacc/particles/regions/RegionBoundsUpdate.hpp:175-214,
 *                               packages/pmacc/include/spmacc/particles/algorithms/FrameIndex.hpp:271-274)
 *   sph_stage_neighbour_counts calculateNeighboursSender (per-particle form of the region CSR)
 *                              (packages/pmacc/include/spmacc/particles/regions/NeighbourRegions.hpp:233-255)
 *   sph_stage_density          UpdateDensity<K>  (include/spearhed/particles/density/DensitySummation.hpp:104-138)
 *   sph_stage_forces           UpdateHydroForces<K> (include/spearhed/sph/HydroForces.hpp:183-214), extended
 *                              with artificial viscosity, SPMHD stress/induction and div-B cleaning
 *   sph_stage_integrate        EulerIntegrate (include/spearhed/particles/pusher/EulerIntegrate.hpp:43-52),
 *                              extended to B and psi, fused with the CFL and conservation reductions
 *   sph_step                   Simulation::runOneStep (include/spearhed/control/Simulation.x.cpp:196-310)
 *   sph_sync                   caravan::syncWait at the end of runOneStep (Simulation.x.cpp:303-307)
 *
 * Conventions: every call returns 0 on success or a negative SPH_E_* code; sph_last_error() gives
 * the message of the last failure on a context. All stage calls are asynchronous on the context's
 * stream until sph_sync(). A context is not thread-safe; distinct contexts may be driven from
 * different host threads. The context owns its device memory; the caller owns host arrays.
 */
#ifndef SPEARHED_SPH_H
#define SPEARHED_SPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------------------------- */
#define SPH_OK 0
#define SPH_E_INVALID (-1)   /* bad argument / configuration */
#define SPH_E_NOMEM (-2)     /* device allocation failed or HBM budget exceeded */
#define SPH_E_DEVICE (-3)    /* HIP runtime error */
#define SPH_E_STATE (-4)     /* call not valid in the context's current state */
#define SPH_E_CAPACITY (-5)  /* a fixed capacity (LDS tile, run table) was exceeded */

/* ---- smoothing kernels (include/spearhed/sph/SphKernel.hpp:45-51 concept) --------------------- */
#define SPH_KERNEL_CUBIC 0     /* M4, support 2h   (include/spearhed/sph/CubicSplineKernel.hpp:40-162) */
#define SPH_KERNEL_QUINTIC 1   /* M5, support 3h   (include/spearhed/sph/QuinticSplineKernel.hpp:39-186) */
#define SPH_KERNEL_WENDLAND 2  /* Wendland C2, support 2h (Wendland 1995; Dehnen & Aly 2012) */

/* ---- precision modes ------------------------------------------------------------------------ */
#define SPH_PREC_FP32 0   /* plain: FP32 absolute positions, FP32 drift, FP32 pair arithmetic */
#define SPH_PREC_MIXED 1  /* FP64 positions/drift/rebase, FP32 pair arithmetic on supercell-relative
                             coordinates, FP64 min-dt and conservation sums */

/* ---- species (include/spearhed/param/speciesDefinition.param:56-88) ------------------------- */
#define SPH_SPECIES_DEFAULT 0   /* Interior, Movable, Source, Thermodynamic: the target of every pass */
#define SPH_SPECIES_BOUNDARY 1  /* Frozen, Source, Thermodynamic: walls, a source only */
#define SPH_SPECIES_TRACER 2    /* Movable only: drifted with its own velocity, never a source */

/* Per-run physics and geometry. One run = one independent simulation box. A configuration is
 * rejected (SPH_E_INVALID, or -1 from the host-only size queries) unless box_lo is finite, box_len,
 * h0, cfl and cfl_force are finite and > 0, gamma is finite and > 1, dt_fixed and cell_factor are
 * finite, and the grid has at most 2^30 supercells. */
typedef struct sph_run_config {
    double box_lo[3];     /* lower corner of the box */
    double box_len[3];    /* edge lengths; periodic axes wrap with this period */
    int32_t periodic[3];  /* 1 = periodic axis (minimum image through the supercell stencil); on a
                             non-periodic axis a position outside the box is binned to the edge
                             supercell (so walls may sit just outside) */
    int32_t kernel;       /* SPH_KERNEL_* */
    float h0;             /* smoothing length that sets the interaction radius H = support * h0
                             (reference: sph.param h0, Simulation.x.cpp:263-264) */
    float gamma;          /* adiabatic index (EquationOfState.hpp:32-47) */
    float alpha_visc;     /* Monaghan artificial viscosity alpha (0 = inviscid, as the reference) */
    float beta_visc;      /* Monaghan artificial viscosity beta; also the signal-velocity coefficient */
    int32_t mhd;          /* 1 = SPMHD (Lorentz force, induction); 0 = hydro only */
    int32_t cleaning;     /* 1 = Tricco & Price (2012) constrained hyperbolic div-B cleaning */
    int32_t borve;        /* 1 = Borve et al. (2001) tensile-instability correction */
    float clean_sigma;    /* cleaning damping: tau = h / (clean_sigma * c_h) */
    float cfl;            /* Courant factor on h / v_sig (reference sph.param cflFactor) */
    float cfl_force;      /* factor on sqrt(h / |a|) */
    double dt_fixed;      /* > 0: constant dt (reference dimension.param dt); <= 0: CFL dt */
    float cell_factor;    /* supercell edge >= cell_factor * H: >= 1 takes the 27-supercell stencil;
                             [1/2, 1) the 125-supercell stencil (5^3 cells of edge >= H/2: about 58% of
                             the candidates, pair_balance 0 only); <= 0 means 1, below 1/2 means 1/2 */
    int32_t reserved;
} sph_run_config;

/* Host-side SoA view of one run, named after the reference attributes
 * (include/spearhed/param/speciesDefinition.param:100-110). Unused pointers may be NULL where
 * noted. Positions are absolute coordinates in FP64 on the host side. */
typedef struct sph_host_soa {
    int64_t n;
    uint64_t* id;          /* particleId */
    uint8_t* mask;         /* multiMask: 1 = live */
    uint8_t* species;      /* SPH_SPECIES_* */
    double* pos[3];        /* position (absolute) */
    float* mass;
    float* vel[3];
    float* density;
    float* smoothing_length;
    float* internal_energy;
    float* bfield[3];      /* magnetic field B (may be NULL for hydro runs: zero) */
    float* psi;            /* cleaning scalar (may be NULL: zero) */
    /* derivative outputs (download only; may be NULL) */
    float* dvdt[3];
    float* dudt;
    float* dbdt[3];
    float* dpsidt;
    float* divb;
    float* vsig;
    uint32_t* n_neighbours;
} sph_host_soa;

/* FP64 global reductions of one run (new in this path: the reference has none on its step). */
typedef struct sph_conserved {
    double mass;
    double momentum[3];
    double energy;   /* sum m (v^2/2 + u + B^2/(2 rho)) */
    double dt_next;  /* CFL dt for the next step: the minimum over live targets; +inf when the run has
                        none (only walls, tracers or dead particles), and a run whose dt is +inf does
                        not advance (no drift, kick or time) until dt_fixed or a host dt is given */
    double time;     /* simulated time of the run */
    double c_h;      /* cleaning speed used in the last force pass */
} sph_conserved;

/* Grid description of one run (queryable for parity checks). */
typedef struct sph_grid_info {
    int32_t n[3];          /* supercells per axis */
    int32_t n_cells;       /* n[0]*n[1]*n[2] */
    double cell[3];        /* supercell edge per axis */
    float radius;          /* interaction radius H */
    int64_t particle_offset; /* first particle of the run in the context's arrays */
    int64_t cell_offset;     /* first supercell of the run in the context's cell table */
} sph_grid_info;

/* Launch tunables chosen by the autotuner (sph_autotune). */
typedef struct sph_tuning {
    int32_t density_chunk;  /* neighbour candidates staged in LDS per chunk by the density pass: 128|256|512 */
    int32_t force_chunk;    /* neighbour records staged in LDS per chunk by the force pass: 64|128 */
    int32_t waves_per_cu;   /* pair passes: 0 = one wave per supercell (default); 1..64 = a persistent
                               grid of waves_per_cu x (compute units) waves, each walking the supercells
                               at a stride of the grid (the reference's GridStride schedule,
                               spmacc/particles/algorithms/FrameSchedule.hpp:92-104, sized as its
                               FixedGrid policy, FrameDispatch.hpp:137-148) */
    int32_t max_lanes_per_target; /* cap on K, the lanes that share one target's candidates: 1..64 */
    int32_t pair_balance;   /* how the accepted pairs of a target are spread over its K lanes:
                               0: each lane its own candidates; 1: the force pass compacts them in
                               LDS and shares them round-robin (the density pass as 0); 2:
                               whole-supercell neighbour lists (density_chunk 512, force_chunk 128) */
} sph_tuning;

typedef struct sph_ctx sph_ctx;

/* ---- context -------------------------------------------------------------------------------- */
/* Create a context holding n_runs independent runs (1 = single run; >1 = ragged batch).
 * capacity[r] = particles of run r. stream: a hipStream_t (NULL = a new stream owned by the ctx).
 * Replaces Simulation::init's buffer set-up (include/spearhed/control/Simulation.x.cpp:312-352). */
int sph_ctx_create(const sph_run_config* runs, const int64_t* capacity, int32_t n_runs,
                   int32_t precision, void* stream, sph_ctx** out);
int sph_ctx_destroy(sph_ctx* ctx);
/* A context with fixed capacities and no runs, whose run set changes between steps (continuous
 * batching for serving): max_cells bounds the summed supercells of resident runs; mhd = 1 allocates
 * the B / psi fields. No counterpart in the reference, which sizes its buffers once at start-up. */
int sph_pool_create(int64_t max_particles, int64_t max_cells, int32_t max_runs, int32_t precision, int32_t mhd,
                    void* stream, sph_ctx** out);
/* New run set = the kept runs (old indices, in the given order; their state and dt/time move with
 * them) followed by the added runs, to be filled with sph_upload or sph_init_run. An added run's
 * first step is its priming step (see sph_upload); kept runs step on undisturbed. Every capacity is
 * checked on the host first (a rejected call changes nothing); the change itself is enqueued on the
 * context's stream (tables through pinned staging, state and per-run scalars moved by kernels), so
 * the host does not wait for the device. */
int sph_pool_set_runs(sph_ctx* ctx, const int32_t* keep, int32_t n_keep, const sph_run_config* add,
                      const int64_t* add_sizes, int32_t n_add);
const char* sph_last_error(const sph_ctx* ctx);
int sph_sync(sph_ctx* ctx);
/* Marks on the context's stream, for hosts that keep work in flight (a serving loop that enqueues
 * the next step before the last one has finished): sph_mark records the current point of the stream
 * and returns its ticket (1, 2, ...); sph_mark_wait blocks until the stream has passed the point of
 * that ticket (every earlier point too), and until the copy-outs (sph_download_async) enqueued before
 * the mark have landed. Waiting on a ticket more than 64 marks old waits for a
 * later point. The analogue of waiting on one step's sender rather than the whole queue. */
int sph_mark(sph_ctx* ctx, uint64_t* ticket);
int sph_mark_wait(sph_ctx* ctx, uint64_t ticket);
int sph_get_grid(sph_ctx* ctx, int32_t run, sph_grid_info* out);
/* Configuration and particle count of run `run` (either pointer may be NULL). */
int sph_get_run(const sph_ctx* ctx, int32_t run, sph_run_config* cfg_out, int64_t* n_out);
/* Runs currently in the context (0 for a null context). */
int32_t sph_run_count(const sph_ctx* ctx);
int sph_set_tuning(sph_ctx* ctx, const sph_tuning* t);
int sph_get_tuning(sph_ctx* ctx, sph_tuning* out);
/* Bytes of HBM a run of n particles needs at a precision (admission by HBM budget). */
int64_t sph_bytes_per_run(const sph_run_config* cfg, int64_t n, int32_t precision);
/* Supercells a run occupies in a pool (its grid or the finest one the autotuner may choose); -1 if
 * the configuration is invalid. Host-only. */
int64_t sph_run_cells(const sph_run_config* cfg);
/* Device bytes sph_pool_create reserves for these capacities; -1 if invalid. Host-only. */
int64_t sph_pool_bytes(int64_t max_particles, int64_t max_cells, int32_t max_runs, int32_t precision, int32_t mhd);

/* ---- initial conditions on the device --------------------------------------------------------
 * A run can be filled on the GPU from lattice blocks instead of uploaded: the analogue of the
 * reference's setup blocks placing particles on a simple cubic lattice
 * (packages/pmacc/include/spmacc/particles/initialization/SC.hpp:117-207: per-axis counts from the
 * block extent and spacing, x truncated with round-off tolerance, other axes rounded; particles at
 * cell centres, x fastest). Each position then gets a uniform jitter of +-jitter * cell edge per
 * axis from Philox4x32-10 (Salmon et al. 2011) keyed by (seed, particle id), so the result does not
 * depend on the launch. The setup fills the fields from the position. */
#define SPH_SETUP_UNIFORM 0     /* density, pressure, velocity, bfield as given */
#define SPH_SETUP_ROTOR 1       /* params: centre x, centre y, radius, edge speed, disc density:
                                   inside the disc rho = disc density, v = rigid rotation */
#define SPH_SETUP_ORSZAG_TANG 2 /* params: v0, B0: v = v0 (-sin 2 pi y, sin 2 pi x, 0),
                                   B = B0 (-sin 2 pi y, sin 4 pi x, 0) */
#define SPH_SETUP_SHEAR_LAYER 3 /* params: dV, density ratio, amplitude, phase: rho * ratio and
                                   vx = +dV/2 for |y - 1/2| < 1/4, else vx = -dV/2;
                                   vy = amplitude sin(4 pi x + phase) */

typedef struct sph_lattice_block {
    double lo[3], hi[3];  /* the block's region */
    double spacing;       /* nominal lattice spacing */
    double density;       /* background density: mass per particle = density * cell volume */
    double pressure;      /* u = pressure / ((gamma - 1) density) */
    float velocity[3];
    float bfield[3];
    int32_t species;      /* SPH_SPECIES_* */
    int32_t setup;        /* SPH_SETUP_* */
    double params[6];
} sph_lattice_block;

/* Particles the blocks place (host-only); -1 if a block is invalid. */
int64_t sph_lattice_sites(const sph_lattice_block* blocks, int32_t n_blocks);
/* Fill run `run` (whose size must equal the blocks' sites) on the device: ids 0..n-1 in block
 * order, smoothing length h0 of the run, then the setup's fields. Like an upload, the run restarts
 * at t = 0 and its next step primes it. Enqueued on the context's stream; the host does not wait. */
int sph_init_run(sph_ctx* ctx, int32_t run, const sph_lattice_block* blocks, int32_t n_blocks, uint64_t seed,
                 double jitter);

/* ---- data movement (the caller owns the host arrays) ---------------------------------------- */
/* New initial data for a run: its time restarts at 0 and its next step is a priming step (density,
 * forces and its first CFL dt, with no drift, no kick and no time advance: the arithmetic of
 * sph_prime for that run alone). Particle ids must be unique within the run. */
int sph_upload(sph_ctx* ctx, int32_t run, const sph_host_soa* src);
/* Downloads in the context's current (supercell-sorted) order; compare by id, never by position. */
int sph_download(sph_ctx* ctx, int32_t run, sph_host_soa* dst);
/* Copy-out for responses (the step after the path: the reference copies selected species fields
 * out for output, spmacc/particles/algorithms/CopyParticlesToDynSoA.hpp, include/spearhed/memory.hpp:42-49):
 * enqueue device-to-host copies of the non-NULL columns of `dst` (and, if `cons` is non-NULL, the
 * run's FP64 sums and next dt as its last step left them) and return. The copies hold the run as it
 * is at this point of the context's stream. In a pool (sph_pool_create) whose snapshot ring (16
 * bytes per particle of capacity, counted in sph_pool_bytes) holds the columns, they are first
 * copied within the device on the context's stream and then go out to the host on a second stream
 * of the library's, overlapping the steps enqueued after them; otherwise they go out on the
 * context's stream. The host buffers are valid after sph_sync, sph_download, sph_upload, or
 * sph_mark_wait on a mark taken after this call; they should be pinned (sph_host_alloc) for the
 * copies to run asynchronously. Positions are available in mixed-precision contexts only. */
int sph_download_async(sph_ctx* ctx, int32_t run, const sph_host_soa* dst, sph_conserved* cons);
/* Pinned host memory for sph_download_async / sph_upload staging. */
int sph_host_alloc(int64_t bytes, void** out);
int sph_host_free(void* p);

/* ---- stages (Simulation::runOneStep order: drift, resort, density, forces, integrate) -------- */
int sph_stage_drift(sph_ctx* ctx, double dt);      /* dt <= 0: use each run's device-resident CFL dt */
int sph_stage_resort(sph_ctx* ctx);
int sph_stage_neighbour_counts(sph_ctx* ctx);      /* fills n_neighbours (download to read) */
int sph_stage_density(sph_ctx* ctx);
int sph_stage_forces(sph_ctx* ctx);
int sph_stage_integrate(sph_ctx* ctx, double dt);  /* dt <= 0: device-resident CFL dt */
/* Reductions over the current state: per-run CFL dt from the last force pass, and FP64 sums. */
int sph_reduce_min_dt(sph_ctx* ctx, double* dt_per_run);
int sph_reduce_conserved(sph_ctx* ctx, sph_conserved* per_run);
/* Resort + density + forces + CFL reduction on the initial state: sets each run's first dt. */
int sph_prime(sph_ctx* ctx);
/* Whole steps: drift, resort, density, forces, integrate (+ reductions), n_steps times. */
int sph_step(sph_ctx* ctx, int32_t n_steps);
/* Capture two whole steps of every run into a hipGraph (a step swaps the double-buffered state,
 * so two return it to where capture began); sph_graph_launch runs n_steps steps by replaying it,
 * taking one step eagerly where the parity or an odd count needs it. Results equal sph_step's.
 * The graph holds the force pass's variant for runs whose particles share one smoothing length; an
 * upload that changes whether every run does so drops the graph (capture again). */
int sph_graph_capture(sph_ctx* ctx);
int sph_graph_launch(sph_ctx* ctx, int32_t n_steps);
/* Per-particle sort results of the last resort (run-local, current order). */
int sph_get_sort(sph_ctx* ctx, int32_t run, uint32_t* cell_key, uint32_t* cell_start);

/* ---- measurement ---------------------------------------------------------------------------- */
/* Per-stage HIP-event timing of the next steps (sph_step), recorded on the context's stream; the
 * analogue of the reference's per-step TimeInterval report (SimulationHelper.cpp:192-195). */
#define SPH_NUM_STAGES 5  /* drift+bin, resort, density, forces, integrate */
typedef struct sph_profile {
    int32_t steps;                    /* steps recorded */
    double stage_ms[SPH_NUM_STAGES];  /* mean per step */
    double step_ms_mean, step_ms_p50, step_ms_p99;
} sph_profile;
int sph_profile_begin(sph_ctx* ctx, int32_t max_steps);
/* Synchronises; step_ms (may be NULL) receives the per-step durations. */
int sph_profile_end(sph_ctx* ctx, sph_profile* out, double* step_ms);

/* A plain device-to-device copy of `bytes` (reps times), timed with HIP events: the known-size
 * reference that HBM counter readings (FETCH_SIZE / WRITE_SIZE) are calibrated against. */
int sph_probe_copy(int64_t bytes, int32_t reps, double* ms_per_copy);

/* ---- density without device allocation -------------------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc) and so cannot create a
 * context: the resort and density pass of sph_stage_resort + sph_stage_density, run on one run whose
 * particles the caller holds in device memory, with the library's own working memory instead of a
 * context's. Results equal sph_upload + sph_stage_density + sph_download of the same particles with
 * ids 0..n-1 in input order (particles of a supercell are ordered by input index), bit for bit.
 *
 * The working memory is one array of at most SPH_NOALLOC_SCRATCH_BYTES per GPU, held from the first
 * call on that GPU until the process ends; the call allocates nothing. Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams or host threads) race and must be
 * serialised by the caller. Sizes beyond SPH_NOALLOC_MAX_PARTICLES particles or
 * SPH_NOALLOC_MAX_CELLS supercells are rejected with SPH_E_CAPACITY before anything is launched. */
#define SPH_NOALLOC_MAX_PARTICLES 262144
#define SPH_NOALLOC_MAX_CELLS 262144
#define SPH_NOALLOC_SCRATCH_BYTES (48ll << 20)

/* Device columns of one run (device pointers, caller-owned, read only). */
typedef struct sph_device_soa {
    int64_t n;
    const uint8_t* mask;          /* multiMask, 1 = live; NULL: every particle live */
    const uint8_t* species;       /* SPH_SPECIES_*; NULL: every particle SPH_SPECIES_DEFAULT */
    const double* pos[3];         /* absolute positions, FP64: required in SPH_PREC_MIXED */
    const float* pos32[3];        /* absolute positions, FP32: required in SPH_PREC_FP32 */
    const float* mass;            /* required */
    const float* smoothing_length; /* required; finite and > 0 (not checked: it stays on the device) */
} sph_device_soa;

/* Density and neighbour counts of the particles in `in`, written to the device arrays density_out
 * and n_neighbours_out (n entries each, input order; 0 for particles that are not live Default
 * particles), asynchronously on `stream` (a hipStream_t; NULL = the null stream). errbuf (may be
 * NULL) receives the message of a failure. The smoothing kernel, radius and box come from cfg; the
 * launch shape is a fresh context's default tuning. */
int sph_density_noalloc(const sph_run_config* cfg, int32_t precision, const sph_device_soa* in,
                        float* density_out, uint32_t* n_neighbours_out, void* stream, char* errbuf,
                        int32_t errbuf_len);

/* ---- autotuner ------------------------------------------------------------------------------ */
/* Time n_steps steps for each candidate tuning on the context's current data (state restored
 * afterwards) and keep the fastest. Writes the chosen tuning to *best if non-NULL. The state is
 * kept in a device copy of the context's memory for the duration of the search (one more
 * allocation of the context's size, freed before returning; SPH_E_DEVICE if it does not fit). */
int sph_autotune(sph_ctx* ctx, int32_t n_steps, sph_tuning* best);

#ifdef __cplusplus
}
#endif
#endif /* SPEARHED_SPH_H */
