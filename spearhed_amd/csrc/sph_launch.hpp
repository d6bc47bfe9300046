// sph_launch.hpp -- host-side launch interface of the gfx950 kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include "sph_device.hpp"

namespace sph {

struct LaunchCtx {
    const RunDev* runs;  // device run table
    int nruns;
    int ncell_tot;
    int nkb;             // particle-kernel blocks (run-aligned, 256 particles each)
    int mixed;
    int has_mhd;
    int density_chunk;
    int force_chunk;
    int kmax;     // max lanes per target in the pair passes
    int balance;  // pair_balance: 1 = the force pass shares a target's accepted pairs among its K lanes,
                  // 2 = whole-supercell neighbour lists
    int kern;     // smoothing kernel of every run (SPH_KERNEL_*), or -1 when runs differ
    int one_h;    // every run's particles share one smoothing length (the run's own)
    int rmax;     // largest stencil reach of the runs: 1 (supercells of edge >= H) or 2 (edge >= H/2; pair
                  // passes take the chunked lane loop, pair_balance 0)
    int persistent_waves;  // pair passes: 0 = one wave per supercell; else a persistent grid of this many
                           // waves (a multiple of 8) walking the supercells at a stride of the grid
    StateBuf S;          // current buffer set
    WorkBuf W;
    hipStream_t stream;
};

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
// Copy the state of n_keep runs from A to B: run k's particles [src, src + count) to [dst, dst + count),
// with (dst, src, count) at tab[3k..3k+2] on the device, dst ascending from 0 and runs contiguous.
// Per-run dt, time and priming flag of a new run set of n_new runs from the n_old old ones (W.run_src:
// the old run each new run continues, or -1 for a fresh run); runs is the new run table.
hipError_t launch_remap_runs(const RunDev* runs, const WorkBuf& W, int n_old, int n_new, hipStream_t s);
// New initial data for one run: time 0, dt its fixed dt or 0, primed by its next step.
hipError_t launch_restart_run(const RunDev* runs, const WorkBuf& W, int run, hipStream_t s);
hipError_t launch_move_runs(const StateBuf& A, const StateBuf& B, const int64_t* tab, int n_keep, int64_t total,
                            int mixed, int mhd, hipStream_t s);
hipError_t launch_drift_bin(const LaunchCtx& c, double dt_host, int do_drift);
hipError_t launch_scan(const LaunchCtx& c);
hipError_t launch_scatter_gather(const LaunchCtx& c, const StateBuf& Snew);
hipError_t launch_density(const LaunchCtx& c);
hipError_t launch_forces(const LaunchCtx& c);
hipError_t launch_kick(const LaunchCtx& c, double dt_host, int do_kick);

}  // namespace sph
