// sph_context.cpp -- host runtime behind the C ABI (include/spearhed_sph.h): device memory layout,
// the stepper that reproduces Simulation::runOneStep's stage order
// (include/spearhed/control/Simulation.x.cpp:196-310), hipGraph capture and the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../../include/spearhed_sph.h"
#include "sph_device.hpp"
#include "sph_launch.hpp"
#include "sph_setup.hpp"

using namespace sph;

struct sph_ctx {
    std::string err;
    int prec = SPH_PREC_MIXED;
    bool mixed = true;
    bool has_mhd = false;
    int nruns = 0;
    std::vector<sph_run_config> cfg;
    std::vector<RunDev> runs_h;
    // per run: every particle has the same smoothing length (set by sph_upload and sph_init_run;
    // the state never changes h), so the force pass may take its one-h instantiation
    std::vector<uint8_t> one_h;
    RunDev* runs_d = nullptr;
    int64_t ntot = 0;
    int ncell_tot = 0;
    int ncell_cap = 0;
    int nkb = 0;
    int64_t cap_particles = 0;
    int cap_runs = 0;
    int cap_kb = 0;
    bool pool_mhd = false;
    StateBuf S[2] = {};
    int cur = 0;
    WorkBuf W = {};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cu = 256;  // compute units of the context's device (sizes the persistent pair-pass grid)
    sph_tuning tune = {128, 64, 0, 64, 0};  // fastest measured on the 2M-particle MHD turbulence step
    void* arena = nullptr;
    size_t arena_bytes = 0;
    // pinned staging of the run tables, two halves used in turn (an event marks when the copies out
    // of a half have run), so that a change of the run set is enqueued without waiting on the host
    char* stage = nullptr;
    size_t stage_half = 0;
    int stage_next = 0;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // marks on the stream (sph_mark / sph_mark_wait): a ring of events, ticket t in slot t % kMarks
    static constexpr int kMarks = 64;
    hipEvent_t marks[kMarks] = {};
    uint64_t next_mark = 1;
    // overlapped copy-out (pools): sph_download_async snapshots a run's columns into a device ring
    // on the context's stream, and a second stream copies them to the host while the next steps
    // run (the next step's re-sort rewrites the state the run was downloaded from)
    struct SnapCopy {
        int64_t off, bytes;
        hipEvent_t done;  // recorded on copy_stream after the copies out of [off, off + bytes)
    };
    char* snap = nullptr;
    int64_t snap_bytes = 0;
    int64_t snap_head = 0;  // end of the newest region in use
    std::deque<SnapCopy> snap_q;  // regions whose copies out may be in flight, oldest first
    std::vector<hipEvent_t> snap_free;
    hipEvent_t snap_ready = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_marks[kMarks] = {};
    bool copy_marked[kMarks] = {};
    bool binned = false;   // keys + histogram valid for the current positions
    bool sorted = false;   // buffers sorted by (cell, id) for the current positions
    bool dens_valid = false;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_parity = -1;
    // profiling: (SPH_NUM_STAGES + 1) events per recorded step
    std::vector<hipEvent_t> prof_ev;
    int prof_cap = 0;
    int prof_steps = 0;
    bool profiling = false;
};

namespace {

int fail(sph_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPTRY(ctx, expr)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), SPH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int support_of(int kernel) { return kernel == SPH_KERNEL_QUINTIC ? 3 : 2; }

float f0_of(int kernel)
{
    if (kernel == SPH_KERNEL_QUINTIC) return 66.0f;  // 3^5 - 6*2^5 + 15
    return 1.0f;
}

float sigma_of(int kernel)
{
    const float pi = 3.14159265358979323846f;
    if (kernel == SPH_KERNEL_CUBIC) return 1.0f / pi;
    if (kernel == SPH_KERNEL_QUINTIC) return 1.0f / (120.0f * pi);
    return 21.0f / (16.0f * pi);
}

// Effective supercell factor: <= 0 (unset) is 1; below 1/2 it is 1/2. Edges of at least H take the
// 27-supercell stencil, edges of at least H/2 the 125-supercell stencil (reach 2).
double cell_factor_of(float cf)
{
    if (!(cf > 0.0f)) return 1.0;
    return cf < 0.5f ? 0.5 : (double)cf;
}

// Supercell grid of one run: per axis the largest count whose edge is >= cell_factor*H*(1+1e-4).
bool make_run(const sph_run_config& c, RunDev& R, std::string& why)
{
    memset(&R, 0, sizeof(R));
    if (!(c.h0 > 0.0f) || !std::isfinite(c.h0)) { why = "h0 must be finite and > 0"; return false; }
    if (c.kernel < 0 || c.kernel > 2) { why = "unknown kernel id"; return false; }
    if (!(c.gamma > 1.0f) || !std::isfinite(c.gamma)) { why = "gamma must be finite and > 1"; return false; }
    if (!(c.cfl > 0.0f) || !std::isfinite(c.cfl) || !(c.cfl_force > 0.0f) || !std::isfinite(c.cfl_force)) {
        why = "cfl and cfl_force must be finite and > 0";
        return false;
    }
    if (std::isnan(c.dt_fixed) || std::isinf(c.dt_fixed)) { why = "dt_fixed must be finite"; return false; }
    if (std::isnan(c.cell_factor) || std::isinf(c.cell_factor)) { why = "cell_factor must be finite"; return false; }
    R.support = support_of(c.kernel);
    R.H = (float)R.support * c.h0;
    R.H2 = R.H * R.H;
    const double f = cell_factor_of(c.cell_factor);
    R.reach = f < 1.0 ? 2 : 1;
    int64_t nc = 1;
    for (int a = 0; a < 3; ++a) {
        const double L = c.box_len[a];
        if (!(L > 0) || !std::isfinite(L) || !std::isfinite(c.box_lo[a])) {
            why = "box_len must be finite and > 0, box_lo finite";
            return false;
        }
        if (c.periodic[a] && L < (double)R.H * (1.0 + 1e-4)) {
            why = "periodic box edge shorter than the interaction radius";
            return false;
        }
        const double nd = std::floor(L / ((double)R.H * f * (1.0 + 1e-4)));
        if (nd > (double)(1 << 30)) { why = "too many supercells"; return false; }
        int n = (int)nd;
        if (n < 1) n = 1;
        R.n[a] = n;
        R.periodic[a] = c.periodic[a] ? 1 : 0;
        R.lo[a] = c.box_lo[a];
        R.L[a] = L;
        R.s[a] = L / (double)n;
        R.inv_s[a] = (double)n / L;
        R.lo32[a] = (float)R.lo[a];
        R.L32[a] = (float)L;
        R.s32[a] = (float)R.s[a];
        R.inv_s32[a] = (float)R.inv_s[a];
        nc *= n;
    }
    if (nc > (1LL << 30)) { why = "too many supercells"; return false; }
    R.ncell = (int32_t)nc;
    R.kernel = c.kernel;
    R.mhd = c.mhd ? 1 : 0;
    R.cleaning = (c.mhd && c.cleaning) ? 1 : 0;
    R.gamma = c.gamma;
    R.alpha = c.alpha_visc;
    R.beta = c.beta_visc;
    R.borve = (c.mhd && c.borve) ? 1.0f : 0.0f;
    R.clean_sigma = c.clean_sigma;
    R.cfl = c.cfl;
    R.cfl_force = c.cfl_force;
    R.sigma = sigma_of(c.kernel);
    R.f0 = f0_of(c.kernel);
    R.dt_fixed = c.dt_fixed;
    return true;
}

struct Carver {
    char* base;
    size_t off = 0;
    bool dry;
    template <class T>
    T* take(int64_t count)
    {
        off = (off + 255) & ~size_t(255);
        T* p = dry ? nullptr : reinterpret_cast<T*>(base + off);
        off += sizeof(T) * (size_t)std::max<int64_t>(count, 1);
        return p;
    }
};

void carve(Carver& cv, sph_ctx* c, int64_t N, int ncell_cap, int nkb, int nruns)
{
    for (int b = 0; b < 2; ++b) {
        StateBuf& S = c->S[b];
        S.id = cv.take<uint64_t>(N);
        S.mask = cv.take<uint8_t>(N);
        S.species = cv.take<uint8_t>(N);
        for (int a = 0; a < 3; ++a) {
            if (c->mixed) { S.pos64[a] = cv.take<double>(N); S.pos32[a] = nullptr; }
            else { S.pos32[a] = cv.take<float>(N); S.pos64[a] = nullptr; }
            S.vel[a] = cv.take<float>(N);
            S.B[a] = c->has_mhd ? cv.take<float>(N) : nullptr;
        }
        S.mass = cv.take<float>(N);
        S.h = cv.take<float>(N);
        S.u = cv.take<float>(N);
        S.rho = cv.take<float>(N);
        S.psi = c->has_mhd ? cv.take<float>(N) : nullptr;
    }
    WorkBuf& W = c->W;
    for (int a = 0; a < 3; ++a) {
        W.dvdt[a] = cv.take<float>(N);
        W.dbdt[a] = c->has_mhd ? cv.take<float>(N) : nullptr;
    }
    W.dudt = cv.take<float>(N);
    W.dpsidt = c->has_mhd ? cv.take<float>(N) : nullptr;
    W.divb = c->has_mhd ? cv.take<float>(N) : nullptr;
    W.vsig = cv.take<float>(N);
    W.nnb = cv.take<uint32_t>(N);
    W.key = cv.take<uint32_t>(N);
    W.rank = cv.take<uint32_t>(N);
    W.tmp = cv.take<uint32_t>(N);
    W.tmpid = cv.take<uint64_t>(N);
    W.nbmask = cv.take<uint32_t>((int64_t)ncell_cap * kMaskWords * 64);
    W.count = cv.take<uint32_t>(ncell_cap);
    W.cs = cv.take<uint32_t>((int64_t)ncell_cap + 1);
    W.bsum = cv.take<uint32_t>((ncell_cap + 2047) / 2048 + 1);
    W.P0 = cv.take<float4>(N);
    W.P1 = cv.take<float4>(N);
    W.P2 = cv.take<float4>(N);
    W.P3 = cv.take<float4>(N);
    W.P4 = c->has_mhd ? cv.take<float4>(N) : nullptr;
    W.dt = cv.take<double>(nruns);
    W.time = cv.take<double>(nruns);
    W.chbits = cv.take<uint32_t>(nruns);
    W.partial = cv.take<double>((int64_t)nkb * 6);
    W.cons = cv.take<double>((int64_t)nruns * 8);
    W.blk_run = cv.take<int32_t>(nkb);
    W.prime_flag = cv.take<int32_t>(nruns);
    W.move_tab = cv.take<int64_t>((int64_t)nruns * 3);
    W.run_src = cv.take<int32_t>(nruns);
    W.run_tmp = cv.take<double>((int64_t)nruns * 3);
    c->snap = c->snap_bytes > 0 ? cv.take<char>(c->snap_bytes) : nullptr;
}

// Device ring of a pool for snapshots of downloaded runs: 16 bytes per particle of capacity, so a
// run of up to a quarter of the capacity with a 64-byte response goes out through it.
int64_t pool_snap_bytes(int64_t max_particles) { return ((max_particles * 16 + 255) / 256) * 256; }

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Bytes of one staging half: run table, block -> run table, run sources, move table.
size_t stage_half_bytes(int cap_r, int cap_kb)
{
    return align256(sizeof(RunDev) * (size_t)cap_r) + align256(4 * (size_t)cap_kb) + align256(4 * (size_t)cap_r) +
           align256(24 * (size_t)cap_r);
}

LaunchCtx lctx(sph_ctx* c)
{
    LaunchCtx L{};
    L.runs = c->runs_d;
    L.nruns = c->nruns;
    L.ncell_tot = c->ncell_tot;
    L.nkb = c->nkb;
    L.mixed = c->mixed ? 1 : 0;
    L.has_mhd = c->has_mhd ? 1 : 0;
    L.density_chunk = c->tune.density_chunk;
    L.kmax = std::max(1, std::min(64, c->tune.max_lanes_per_target));
    L.S = c->S[c->cur];
    L.force_chunk = c->tune.force_chunk;
    L.balance = c->tune.pair_balance < 0 ? 0 : (c->tune.pair_balance > 2 ? 2 : c->tune.pair_balance);
    L.one_h = c->nruns > 0 ? 1 : 0;
    for (int r = 0; r < c->nruns; ++r) L.one_h = L.one_h && c->one_h[r];
    L.kern = c->nruns > 0 ? c->cfg[0].kernel : -1;
    for (int r = 1; r < c->nruns; ++r)
        if (c->cfg[r].kernel != L.kern) L.kern = -1;
    L.rmax = 1;
    for (const RunDev& R : c->runs_h) L.rmax = std::max(L.rmax, (int)R.reach);
    L.persistent_waves = 0;
    if (c->tune.waves_per_cu > 0) {
        const int64_t w = (int64_t)c->tune.waves_per_cu * c->num_cu;
        L.persistent_waves = (int)std::max<int64_t>(8, w / 8 * 8);
    }
    L.W = c->W;
    L.stream = c->stream;
    return L;
}

void invalidate_graph(sph_ctx* c)
{
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_parity = -1;
}

// A run's smoothing-length flag after an upload or fill. The force pass's kernel variant follows
// the flags of all runs; a captured graph holds the variant it was captured with, so it is dropped
// when that changes (sph_graph_launch then reports that no graph is captured).
void set_one_h(sph_ctx* c, int run, bool one)
{
    bool before = c->nruns > 0, after = c->nruns > 0;
    for (int r = 0; r < c->nruns; ++r) {
        before = before && c->one_h[r];
        after = after && (r == run ? one : c->one_h[r] != 0);
    }
    c->one_h[run] = one ? 1 : 0;
    if (before != after) invalidate_graph(c);
}

int do_bin(sph_ctx* c, double dt, int drift)
{
    if (c->nruns == 0) return SPH_OK;  // an empty pool: nothing to launch
    if (c->binned) HIPTRY(c, launch_fill_u32(c->W.count, 0u, c->ncell_tot, c->stream));
    LaunchCtx L = lctx(c);
    HIPTRY(c, launch_drift_bin(L, dt, drift));
    c->binned = true;
    c->sorted = false;
    c->dens_valid = false;
    return SPH_OK;
}

int do_resort(sph_ctx* c)
{
    if (c->nruns == 0) return SPH_OK;  // an empty pool: nothing to launch
    if (!c->binned) {
        int rc = do_bin(c, 0.0, 0);
        if (rc) return rc;
    }
    LaunchCtx L = lctx(c);
    HIPTRY(c, launch_scan(L));
    HIPTRY(c, launch_scatter_gather(L, c->S[c->cur ^ 1]));
    c->cur ^= 1;
    c->binned = false;
    c->sorted = true;
    c->dens_valid = false;
    return SPH_OK;
}

int do_density(sph_ctx* c)
{
    if (c->nruns == 0) return SPH_OK;  // an empty pool: nothing to launch
    if (!c->sorted) {
        int rc = do_resort(c);
        if (rc) return rc;
    }
    LaunchCtx L = lctx(c);
    HIPTRY(c, launch_density(L));
    c->dens_valid = true;
    return SPH_OK;
}

int do_forces(sph_ctx* c)
{
    if (c->nruns == 0) return SPH_OK;  // an empty pool: nothing to launch
    if (!c->dens_valid) return fail(c, SPH_E_STATE, "force pass needs a density pass after the last resort");
    LaunchCtx L = lctx(c);
    HIPTRY(c, launch_forces(L));
    return SPH_OK;
}

int do_kick(sph_ctx* c, double dt, int kick)
{
    if (c->nruns == 0) return SPH_OK;  // an empty pool: nothing to launch
    LaunchCtx L = lctx(c);
    HIPTRY(c, launch_kick(L, dt, kick));
    return SPH_OK;
}

int prof_mark(sph_ctx* c, int k)
{
    if (!c->profiling || c->prof_steps >= c->prof_cap) return SPH_OK;
    HIPTRY(c, hipEventRecord(c->prof_ev[(size_t)c->prof_steps * (SPH_NUM_STAGES + 1) + k], c->stream));
    return SPH_OK;
}

// One whole step in runOneStep order: drift (+bin), resort, density, forces, kick (+reductions).
int do_step(sph_ctx* c)
{
    int rc = prof_mark(c, 0);
    if (!rc) rc = do_bin(c, 0.0, 1);
    if (!rc) rc = prof_mark(c, 1);
    if (!rc) rc = do_resort(c);
    if (!rc) rc = prof_mark(c, 2);
    if (!rc) rc = do_density(c);
    if (!rc) rc = prof_mark(c, 3);
    if (!rc) rc = do_forces(c);
    if (!rc) rc = prof_mark(c, 4);
    if (!rc) rc = do_kick(c, 0.0, 1);
    if (!rc) rc = prof_mark(c, 5);
    if (!rc && c->profiling && c->prof_steps < c->prof_cap) ++c->prof_steps;
    return rc;
}

// Allocate the arena for fixed capacities (particles, supercells, runs, particle-kernel blocks).
int reserve(sph_ctx* c, int64_t cap_p, int64_t cap_c, int cap_r, int cap_kb)
{
    if (cap_p < 1 || cap_r < 1 || cap_c < 1) return fail(c, SPH_E_INVALID, "capacities must be >= 1");
    // particle indices and the cell-count scan are 32-bit signed on the device
    if (cap_p > (1LL << 30)) return fail(c, SPH_E_CAPACITY, "more than 2^30 particles in one context");
    if (cap_c >= (1LL << 30)) return fail(c, SPH_E_CAPACITY, "too many supercells");
    c->cap_particles = cap_p;
    c->ncell_cap = (int)cap_c;
    c->cap_runs = cap_r;
    c->cap_kb = cap_kb;
    Carver dry{nullptr, 0, true};
    carve(dry, c, cap_p, c->ncell_cap, cap_kb, cap_r);
    c->arena_bytes = ((dry.off + 255) & ~size_t(255)) + 256;
    HIPTRY(c, hipMalloc(&c->arena, c->arena_bytes));
    Carver real{reinterpret_cast<char*>(c->arena), 0, false};
    carve(real, c, cap_p, c->ncell_cap, cap_kb, cap_r);
    // Cleared on the context's stream: the tables that layout() uploads next go through that
    // (non-blocking) stream, which is not ordered after work on the null stream.
    HIPTRY(c, launch_fill_u32(reinterpret_cast<uint32_t*>(c->arena), 0u, (int64_t)(c->arena_bytes / 4), c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    HIPTRY(c, hipMalloc(&c->runs_d, sizeof(RunDev) * cap_r));
    c->stage_half = stage_half_bytes(cap_r, cap_kb);
    HIPTRY(c, hipHostMalloc((void**)&c->stage, 2 * c->stage_half, hipHostMallocDefault));
    for (auto& e : c->stage_ev) {
        HIPTRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPTRY(c, hipEventRecord(e, c->stream));
    }
    return SPH_OK;
}

// Upload the run table (and, when given, the block -> run table, the run sources and the move table)
// through the next staging half, on the context's stream. The host waits only until the copies out
// of that half from two uploads ago have run.
int stage_tables(sph_ctx* c, const std::vector<RunDev>& runs, int kb, const std::vector<int32_t>* src,
                 const std::vector<int64_t>* move)
{
    const int h = c->stage_next;
    c->stage_next ^= 1;
    HIPTRY(c, hipEventSynchronize(c->stage_ev[h]));
    char* base = c->stage + (size_t)h * c->stage_half;
    const size_t cap_r = (size_t)c->cap_runs, cap_kb = (size_t)c->cap_kb;
    RunDev* t_runs = reinterpret_cast<RunDev*>(base);
    int32_t* t_blk = reinterpret_cast<int32_t*>(base + align256(sizeof(RunDev) * cap_r));
    int32_t* t_src = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(t_blk) + align256(4 * cap_kb));
    int64_t* t_move = reinterpret_cast<int64_t*>(reinterpret_cast<char*>(t_src) + align256(4 * cap_r));
    const size_t nr = runs.size();
    if (nr > 0) {
        std::copy(runs.begin(), runs.end(), t_runs);
        HIPTRY(c, hipMemcpyAsync(c->runs_d, t_runs, sizeof(RunDev) * nr, hipMemcpyHostToDevice, c->stream));
    }
    if (src && nr > 0) {
        for (size_t r = 0; r < nr; ++r)
            for (int b = 0; b < runs[r].kb_count; ++b) t_blk[runs[r].kb_first + b] = (int32_t)r;
        HIPTRY(c, hipMemcpyAsync(c->W.blk_run, t_blk, sizeof(int32_t) * kb, hipMemcpyHostToDevice, c->stream));
        std::copy(src->begin(), src->end(), t_src);
        HIPTRY(c, hipMemcpyAsync(c->W.run_src, t_src, sizeof(int32_t) * nr, hipMemcpyHostToDevice, c->stream));
    }
    if (move && !move->empty()) {
        std::copy(move->begin(), move->end(), t_move);
        HIPTRY(c, hipMemcpyAsync(c->W.move_tab, t_move, sizeof(int64_t) * move->size(), hipMemcpyHostToDevice,
                                 c->stream));
    }
    HIPTRY(c, hipEventRecord(c->stage_ev[h], c->stream));
    return SPH_OK;
}

// Lay runs out contiguously (particles, supercells, kick blocks) and enqueue the run and block tables:
// new run r continues old run src[r] (its dt, time and priming flag) or, for src[r] < 0, starts fresh
// and is primed by its next step. Nothing waits on the device.
// Supercells finer than H are swept by the chunked lane loop only (pair_balance 0).
bool balance_allowed(const std::vector<sph_run_config>& cfg, int pair_balance)
{
    if (pair_balance == 0) return true;
    for (const sph_run_config& r : cfg)
        if (cell_factor_of(r.cell_factor) < 1.0) return false;
    return true;
}
const char* kFineBalance = "supercells finer than H (cell_factor < 1) need pair_balance 0";

int layout(sph_ctx* c, const std::vector<sph_run_config>& cfg, const std::vector<int64_t>& sizes,
           const std::vector<int32_t>& src, const std::vector<int64_t>* move = nullptr)
{
    const int nr = (int)cfg.size();
    if (nr > c->cap_runs) return fail(c, SPH_E_CAPACITY, "run table capacity exceeded");
    if (!balance_allowed(cfg, c->tune.pair_balance)) return fail(c, SPH_E_INVALID, kFineBalance);
    std::vector<RunDev> runs(nr);
    int64_t p0 = 0;
    int64_t c0 = 0;
    int kb = 0;
    for (int r = 0; r < nr; ++r) {
        std::string why;
        if (!make_run(cfg[r], runs[r], why)) return fail(c, SPH_E_INVALID, "run " + std::to_string(r) + ": " + why);
        if (sizes[r] < 1) return fail(c, SPH_E_INVALID, "run size must be >= 1");
        RunDev& R = runs[r];
        R.p0 = p0;
        R.np = sizes[r];
        R.c0 = (int32_t)c0;
        R.kb_first = kb;
        R.kb_count = (int32_t)((R.np + 255) / 256);
        p0 += R.np;
        c0 += R.ncell;
        kb += R.kb_count;
    }
    if (p0 > c->cap_particles) return fail(c, SPH_E_CAPACITY, "particle capacity exceeded");
    if (c0 > c->ncell_cap) return fail(c, SPH_E_CAPACITY, "supercell capacity exceeded");
    if (kb > c->cap_kb) return fail(c, SPH_E_CAPACITY, "block table capacity exceeded");
    const int n_old = c->nruns;
    std::vector<uint8_t> one_h(nr, 0);  // kept runs keep their flag; new runs have no data yet
    for (int r = 0; r < nr; ++r)
        if (src[r] >= 0 && src[r] < (int)c->one_h.size()) one_h[r] = c->one_h[src[r]];
    c->one_h = one_h;
    c->nruns = nr;
    c->cfg = cfg;
    c->runs_h = runs;
    c->ntot = p0;
    c->ncell_tot = (int)c0;
    c->nkb = kb;
    int rc = stage_tables(c, runs, kb, &src, move);
    if (rc) return rc;
    HIPTRY(c, launch_remap_runs(c->runs_d, c->W, n_old, nr, c->stream));
    HIPTRY(c, launch_fill_u32(c->W.count, 0u, c->ncell_cap, c->stream));
    c->binned = false;
    c->sorted = false;
    c->dens_valid = false;
    invalidate_graph(c);
    return SPH_OK;
}

// Re-derive every run's grid (after a cell_factor change) keeping particle offsets.
int regrid(sph_ctx* c)
{
    if (!balance_allowed(c->cfg, c->tune.pair_balance)) return fail(c, SPH_E_INVALID, kFineBalance);
    std::vector<RunDev> runs(c->runs_h);  // committed only if every run fits
    int64_t c0 = 0;
    for (int r = 0; r < c->nruns; ++r) {
        RunDev R;
        std::string why;
        if (!make_run(c->cfg[r], R, why)) return fail(c, SPH_E_INVALID, why);
        R.p0 = runs[r].p0;
        R.np = runs[r].np;
        R.kb_first = runs[r].kb_first;
        R.kb_count = runs[r].kb_count;
        R.c0 = (int32_t)c0;
        c0 += R.ncell;
        runs[r] = R;
        if (c0 > c->ncell_cap) return fail(c, SPH_E_CAPACITY, "supercell table capacity exceeded");
    }
    c->runs_h = runs;
    c->ncell_tot = (int)c0;
    int rc = stage_tables(c, c->runs_h, c->nkb, nullptr, nullptr);
    if (rc) return rc;
    HIPTRY(c, launch_fill_u32(c->W.count, 0u, c->ncell_cap, c->stream));
    c->binned = false;
    c->sorted = false;
    c->dens_valid = false;
    invalidate_graph(c);
    return SPH_OK;
}

template <class T>
int copy_col(sph_ctx* c, T* dst_dev, const T* src_host, int64_t n, T fill)
{
    if (src_host) {
        HIPTRY(c, hipMemcpyAsync(dst_dev, src_host, sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
    } else {
        std::vector<T> z(n, fill);
        HIPTRY(c, hipMemcpyAsync(dst_dev, z.data(), sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
        HIPTRY(c, hipStreamSynchronize(c->stream));
    }
    return SPH_OK;
}

template <class T>
int fetch_col(sph_ctx* c, T* dst_host, const T* src_dev, int64_t n)
{
    if (!dst_host || !src_dev) return SPH_OK;
    HIPTRY(c, hipMemcpyAsync(dst_host, src_dev, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
    return SPH_OK;
}

// Wait for every copy out of the snapshot ring (host-side) and empty the ring.
int join_copies(sph_ctx* c)
{
    if (!c->copy_stream) return SPH_OK;
    HIPTRY(c, hipStreamSynchronize(c->copy_stream));
    for (const auto& e : c->snap_q) c->snap_free.push_back(e.done);
    c->snap_q.clear();
    c->snap_head = 0;
    return SPH_OK;
}

// A region of `need` bytes in the snapshot ring, at `off`. Regions whose copies out have finished
// are taken back first; if that is not enough, the oldest regions are taken back with the context's
// stream waiting (on the device) for their copies out. Returns false if the ring is too small.
bool snap_alloc(sph_ctx* c, int64_t need, int64_t& off, int& rc)
{
    rc = SPH_OK;
    if (need > c->snap_bytes) return false;
    while (!c->snap_q.empty() && hipEventQuery(c->snap_q.front().done) == hipSuccess) {
        c->snap_free.push_back(c->snap_q.front().done);
        c->snap_q.pop_front();
    }
    for (;;) {
        if (c->snap_q.empty()) {
            off = 0;
            break;
        }
        const int64_t tail = c->snap_q.front().off;  // start of the oldest region in use
        if (c->snap_head > tail) {                     // in use: [tail, head)
            if (c->snap_bytes - c->snap_head >= need) { off = c->snap_head; break; }
            if (tail >= need) { off = 0; break; }
        } else if (tail - c->snap_head >= need) {      // in use: [tail, end) and [0, head)
            off = c->snap_head;
            break;
        }
        const hipError_t e = hipStreamWaitEvent(c->stream, c->snap_q.front().done, 0);
        if (e != hipSuccess) {
            rc = fail(c, SPH_E_DEVICE, std::string("hipStreamWaitEvent: ") + hipGetErrorString(e));
            return false;
        }
        c->snap_free.push_back(c->snap_q.front().done);
        c->snap_q.pop_front();
    }
    c->snap_head = off + need;
    return true;
}

// Supercells a run may need: its own grid or the finest (cell_factor 1) the autotuner may pick.
int64_t cells_needed(const sph_run_config& cfg, std::string& why)
{
    RunDev R, Rf;
    if (!make_run(cfg, R, why)) return -1;
    sph_run_config fine = cfg;
    fine.cell_factor = 1.0f;
    make_run(fine, Rf, why);
    return std::max<int64_t>(R.ncell, Rf.ncell);
}

int open_stream(sph_ctx* c, void* stream)
{
    int dev = 0;
    HIPTRY(c, hipGetDevice(&dev));
    HIPTRY(c, hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    if (c->num_cu < 1) c->num_cu = 1;
    if (stream) {
        c->stream = (hipStream_t)stream;
        return SPH_OK;
    }
    HIPTRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
    return SPH_OK;
}

}  // namespace

namespace sph {
// The grid and constants of one run from its configuration (for entry points without a context).
bool make_run_config(const sph_run_config& c, RunDev& R, char* why, int why_len)
{
    std::string w;
    const bool ok = make_run(c, R, w);
    if (!ok && why && why_len > 0) snprintf(why, (size_t)why_len, "%s", w.c_str());
    return ok;
}
}  // namespace sph

// =================================================================================================
// C ABI
// =================================================================================================
extern "C" {

int sph_ctx_create(const sph_run_config* runs, const int64_t* capacity, int32_t n_runs, int32_t precision,
                   void* stream, sph_ctx** out)
{
    if (!out || !runs || !capacity || n_runs < 1) return SPH_E_INVALID;
    *out = nullptr;
    if (precision != SPH_PREC_FP32 && precision != SPH_PREC_MIXED) return SPH_E_INVALID;
    sph_ctx* c = new sph_ctx();
    *out = c;  // returned even on failure so that sph_last_error can report; caller destroys
    c->prec = precision;
    c->mixed = precision == SPH_PREC_MIXED;
    std::vector<sph_run_config> cfg(runs, runs + n_runs);
    std::vector<int64_t> sizes(capacity, capacity + n_runs);
    int64_t cap_p = 0, cap_c = 0;
    int cap_kb = 0;
    for (int r = 0; r < n_runs; ++r) {
        std::string why;
        const int64_t nc = cells_needed(cfg[r], why);
        if (nc < 0) return fail(c, SPH_E_INVALID, "run " + std::to_string(r) + ": " + why);
        if (sizes[r] < 1) return fail(c, SPH_E_INVALID, "run capacity must be >= 1");
        if (sizes[r] > (1LL << 30) || cap_p + sizes[r] > (1LL << 30))
            return fail(c, SPH_E_CAPACITY, "more than 2^30 particles in one context");
        cap_p += sizes[r];
        cap_c += nc;
        cap_kb += (int)((sizes[r] + 255) / 256);
        c->has_mhd = c->has_mhd || cfg[r].mhd;
    }
    int rc = open_stream(c, stream);
    if (!rc) rc = reserve(c, cap_p, cap_c, n_runs, cap_kb);
    if (rc) return rc;
    return layout(c, cfg, sizes, std::vector<int32_t>(n_runs, -1));
}

int sph_pool_create(int64_t max_particles, int64_t max_cells, int32_t max_runs, int32_t precision, int32_t mhd,
                    void* stream, sph_ctx** out)
{
    if (!out || max_particles < 1 || max_cells < 1 || max_runs < 1) return SPH_E_INVALID;
    *out = nullptr;
    if (precision != SPH_PREC_FP32 && precision != SPH_PREC_MIXED) return SPH_E_INVALID;
    sph_ctx* c = new sph_ctx();
    *out = c;
    c->prec = precision;
    c->mixed = precision == SPH_PREC_MIXED;
    c->has_mhd = mhd != 0;
    const int64_t cap_kb = max_particles / 256 + max_runs;
    if (cap_kb >= (1LL << 31)) return fail(c, SPH_E_CAPACITY, "too many particle blocks");
    c->snap_bytes = pool_snap_bytes(max_particles);
    int rc = open_stream(c, stream);
    if (!rc) rc = reserve(c, max_particles, max_cells, max_runs, (int)cap_kb);
    return rc;
}

int sph_pool_set_runs(sph_ctx* c, const int32_t* keep, int32_t n_keep, const sph_run_config* add,
                      const int64_t* add_sizes, int32_t n_add)
{
    if (!c || n_keep < 0 || n_add < 0 || (n_keep && !keep) || (n_add && (!add || !add_sizes))) return SPH_E_INVALID;
    std::vector<char> seen(c->nruns, 0);
    for (int k = 0; k < n_keep; ++k) {
        if (keep[k] < 0 || keep[k] >= c->nruns || seen[keep[k]])
            return fail(c, SPH_E_INVALID, "keep list must name distinct existing runs");
        seen[keep[k]] = 1;
    }
    std::vector<sph_run_config> cfg;
    std::vector<int64_t> sizes;
    for (int k = 0; k < n_keep; ++k) {
        cfg.push_back(c->cfg[keep[k]]);
        sizes.push_back(c->runs_h[keep[k]].np);
    }
    int64_t cells = 0, parts = 0;
    for (int k = 0; k < n_keep; ++k) {
        std::string why;
        cells += cells_needed(c->cfg[keep[k]], why);
        parts += c->runs_h[keep[k]].np;
    }
    for (int a = 0; a < n_add; ++a) {
        std::string why;
        const int64_t nc = cells_needed(add[a], why);
        if (nc < 0) return fail(c, SPH_E_INVALID, "added run " + std::to_string(a) + ": " + why);
        if (add[a].mhd && !c->has_mhd) return fail(c, SPH_E_INVALID, "pool was created without MHD fields");
        if (add_sizes[a] < 1) return fail(c, SPH_E_INVALID, "run size must be >= 1");
        cfg.push_back(add[a]);
        sizes.push_back(add_sizes[a]);
        cells += nc;
        parts += add_sizes[a];
    }
    int64_t blocks = 0;  // particle-kernel blocks of the new run set (each run starts a block)
    for (int64_t sz : sizes) blocks += (sz + 255) / 256;
    // every capacity is checked before anything moves: a rejected call leaves the pool as it was
    if (parts > c->cap_particles || cells > c->ncell_cap || (int)cfg.size() > c->cap_runs || blocks > c->cap_kb)
        return fail(c, SPH_E_CAPACITY, "runs do not fit the pool's capacity");
    if (!balance_allowed(cfg, c->tune.pair_balance)) return fail(c, SPH_E_INVALID, kFineBalance);
    // new run k < n_keep continues old run keep[k] (its dt, time and priming flag move with it on the
    // device: a kept run not yet stepped still primes on its first step); added runs start fresh
    std::vector<int32_t> src;
    for (int k = 0; k < n_keep; ++k) src.push_back(keep[k]);
    for (int a = 0; a < n_add; ++a) src.push_back(-1);
    // the kept runs' state moves into the other buffer set at their new offsets: one table and one
    // kernel for all runs and columns
    std::vector<int64_t> tab;
    int64_t dst = 0;
    for (int k = 0; k < n_keep; ++k) {
        tab.push_back(dst);
        tab.push_back(c->runs_h[keep[k]].p0);
        tab.push_back(c->runs_h[keep[k]].np);
        dst += c->runs_h[keep[k]].np;
    }
    const int from = c->cur;
    int rc = layout(c, cfg, sizes, src, &tab);  // enqueues the tables, including the move table
    if (rc) return rc;
    if (n_keep > 0)
        HIPTRY(c, launch_move_runs(c->S[from], c->S[from ^ 1], c->W.move_tab, n_keep, dst, c->mixed ? 1 : 0,
                                   c->has_mhd ? 1 : 0, c->stream));
    c->cur = from ^ 1;
    return SPH_OK;
}

int sph_ctx_destroy(sph_ctx* c)
{
    if (!c) return SPH_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->copy_stream) hipStreamSynchronize(c->copy_stream);
    invalidate_graph(c);
    for (const auto& e : c->snap_q) hipEventDestroy(e.done);
    for (auto e : c->snap_free) hipEventDestroy(e);
    for (auto e : c->copy_marks)
        if (e) hipEventDestroy(e);
    if (c->snap_ready) hipEventDestroy(c->snap_ready);
    if (c->copy_stream) hipStreamDestroy(c->copy_stream);
    for (auto e : c->prof_ev) hipEventDestroy(e);
    for (auto e : c->stage_ev)
        if (e) hipEventDestroy(e);
    for (auto e : c->marks)
        if (e) hipEventDestroy(e);
    if (c->stage) hipHostFree(c->stage);
    if (c->arena) hipFree(c->arena);
    if (c->runs_d) hipFree(c->runs_d);
    if (c->own_stream) hipStreamDestroy(c->stream);
    delete c;
    return SPH_OK;
}

const char* sph_last_error(const sph_ctx* c) { return c ? c->err.c_str() : "null context"; }

int sph_mark(sph_ctx* c, uint64_t* ticket)
{
    if (!c || !ticket) return SPH_E_INVALID;
    const uint64_t t = c->next_mark;
    const int slot = (int)(t % sph_ctx::kMarks);
    hipEvent_t& e = c->marks[slot];
    if (!e) HIPTRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPTRY(c, hipEventRecord(e, c->stream));
    // and the point of the copy-out stream: copies out enqueued before the mark have landed too
    c->copy_marked[slot] = c->copy_stream != nullptr;
    if (c->copy_stream) {
        hipEvent_t& ce = c->copy_marks[slot];
        if (!ce) HIPTRY(c, hipEventCreateWithFlags(&ce, hipEventDisableTiming));
        HIPTRY(c, hipEventRecord(ce, c->copy_stream));
    }
    ++c->next_mark;
    *ticket = t;
    return SPH_OK;
}

int sph_mark_wait(sph_ctx* c, uint64_t ticket)
{
    if (!c || ticket == 0 || ticket >= c->next_mark) return SPH_E_INVALID;
    // a slot reused by a later mark waits for that later point: the stream has then passed both
    const int slot = (int)(ticket % sph_ctx::kMarks);
    HIPTRY(c, hipEventSynchronize(c->marks[slot]));
    if (c->copy_marked[slot]) HIPTRY(c, hipEventSynchronize(c->copy_marks[slot]));
    return SPH_OK;
}

int sph_sync(sph_ctx* c)
{
    if (!c) return SPH_E_INVALID;
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return join_copies(c);
}

int sph_get_grid(sph_ctx* c, int32_t run, sph_grid_info* o)
{
    if (!c || !o || run < 0 || run >= c->nruns) return SPH_E_INVALID;
    const RunDev& R = c->runs_h[run];
    for (int a = 0; a < 3; ++a) {
        o->n[a] = R.n[a];
        o->cell[a] = R.s[a];
    }
    o->n_cells = R.ncell;
    o->radius = R.H;
    o->particle_offset = R.p0;
    o->cell_offset = R.c0;
    return SPH_OK;
}

int sph_get_run(const sph_ctx* c, int32_t run, sph_run_config* cfg, int64_t* n)
{
    if (!c || run < 0 || run >= c->nruns) return SPH_E_INVALID;
    if (cfg) *cfg = c->cfg[run];
    if (n) *n = c->runs_h[run].np;
    return SPH_OK;
}

int32_t sph_run_count(const sph_ctx* c) { return c ? c->nruns : 0; }

int sph_set_tuning(sph_ctx* c, const sph_tuning* t)
{
    if (!c || !t) return SPH_E_INVALID;
    const int dc = t->density_chunk, fc = t->force_chunk;
    if (t->pair_balance < 0 || t->pair_balance > 2) return fail(c, SPH_E_INVALID, "pair_balance must be 0, 1 or 2");
    if (t->pair_balance == 2) {
        if (dc != 512) return fail(c, SPH_E_INVALID, "whole-cell density_chunk must be 512");
        if (fc != 128) return fail(c, SPH_E_INVALID, "whole-cell force_chunk must be 128");
    } else {
        if (dc != 128 && dc != 256) return fail(c, SPH_E_INVALID, "density_chunk must be 128 or 256");
        if (fc != 64 && fc != 128) return fail(c, SPH_E_INVALID, "force_chunk must be 64 or 128");
    }
    if (t->max_lanes_per_target < 1 || t->max_lanes_per_target > 64)
        return fail(c, SPH_E_INVALID, "max_lanes_per_target must be in [1, 64]");
    if (t->waves_per_cu != 0 && (t->waves_per_cu < 1 || t->waves_per_cu > 64))
        return fail(c, SPH_E_INVALID, "waves_per_cu must be 0 (one wave per supercell) or in [1, 64]");
    if (!balance_allowed(c->cfg, t->pair_balance)) return fail(c, SPH_E_INVALID, kFineBalance);
    c->tune = *t;
    c->dens_valid = false;  // the force pass replays masks laid out for the density pass's tuning
    invalidate_graph(c);
    return SPH_OK;
}

int sph_get_tuning(sph_ctx* c, sph_tuning* o)
{
    if (!c || !o) return SPH_E_INVALID;
    *o = c->tune;
    return SPH_OK;
}

int64_t sph_bytes_per_run(const sph_run_config* cfg, int64_t n, int32_t precision)
{
    if (!cfg || n < 0) return -1;
    RunDev R;
    std::string why;
    if (!make_run(*cfg, R, why)) return -1;  // the configuration as given must be valid
    const int64_t ncell = cells_needed(*cfg, why);
    const int64_t pos = precision == SPH_PREC_MIXED ? 24 : 12;
    const int64_t mhd = cfg->mhd ? 1 : 0;
    const int64_t state = 8 + 2 + pos + 12 + 16 + mhd * 16;
    // work arrays, and the pool's copy-out snapshot ring (16 B per particle of capacity)
    const int64_t work = 12 + 4 + 4 + 4 + 12 + 8 + 64 + mhd * (12 + 4 + 4 + 16) + 16;
    const int64_t blocks = (n + 255) / 256;  // FP64 partial sums (6 doubles) + block->run entry per block
    // dt, time, cleaning speed, sums, flag, move entry, source index, saved scalars
    const int64_t per_run = 8 + 8 + 4 + 8 * 8 + 4 + 3 * 8 + 4 + 3 * 8;
    return n * (2 * state + work) + blocks * (6 * 8 + 4) + ncell * (12 + 4 * kMaskWords * 64) + per_run + 4096;
}

int64_t sph_pool_bytes(int64_t max_particles, int64_t max_cells, int32_t max_runs, int32_t precision, int32_t mhd)
{
    if (max_particles < 1 || max_cells < 1 || max_runs < 1 || max_cells >= (1LL << 30)) return -1;
    sph_ctx tmp;
    tmp.prec = precision;
    tmp.mixed = precision == SPH_PREC_MIXED;
    tmp.has_mhd = mhd != 0;
    tmp.snap_bytes = pool_snap_bytes(max_particles);
    Carver dry{nullptr, 0, true};
    carve(dry, &tmp, max_particles, (int)max_cells, (int)(max_particles / 256 + max_runs), max_runs);
    return (int64_t)dry.off + 256 + (int64_t)sizeof(RunDev) * max_runs;
}

int64_t sph_run_cells(const sph_run_config* cfg)
{
    if (!cfg) return -1;
    std::string why;
    return cells_needed(*cfg, why);
}

int sph_upload(sph_ctx* c, int32_t run, const sph_host_soa* s)
{
    if (!c || !s || run < 0 || run >= c->nruns) return SPH_E_INVALID;
    const RunDev& R = c->runs_h[run];
    if (s->n != R.np) return fail(c, SPH_E_INVALID, "upload: n must equal the run's capacity");
    if (!s->id || !s->pos[0] || !s->pos[1] || !s->pos[2] || !s->mass || !s->vel[0] || !s->vel[1] || !s->vel[2] ||
        !s->smoothing_length || !s->internal_energy)
        return fail(c, SPH_E_INVALID, "upload: id, pos, mass, vel, smoothing_length, internal_energy are required");
    {
        // ids must be unique within a run (the re-sort ranks particles of a supercell by id):
        // a bitmap when the ids are dense (the usual 0..n-1), a sort otherwise
        const uint64_t* id = s->id;
        const int64_t n = s->n;
        const uint64_t mx = n ? *std::max_element(id, id + n) : 0;
        bool dup = false;
        if (mx < (uint64_t)(4 * n + 64)) {
            std::vector<uint64_t> seen((size_t)(mx / 64 + 1), 0);
            for (int64_t i = 0; i < n && !dup; ++i) {
                const uint64_t bit = 1ull << (id[i] & 63);
                uint64_t& w = seen[(size_t)(id[i] >> 6)];
                dup = (w & bit) != 0;
                w |= bit;
            }
        } else {
            std::vector<uint64_t> ids(id, id + n);
            std::sort(ids.begin(), ids.end());
            dup = std::adjacent_find(ids.begin(), ids.end()) != ids.end();
        }
        if (dup) return fail(c, SPH_E_INVALID, "upload: particle ids must be unique within a run");
    }
    bool one_h = true;
    for (int64_t i = 0; i < s->n; ++i) {
        if (!(s->smoothing_length[i] > 0.0f) || !std::isfinite(s->smoothing_length[i]))
            return fail(c, SPH_E_INVALID, "upload: smoothing_length must be finite and > 0");
        one_h = one_h && s->smoothing_length[i] == s->smoothing_length[0];
    }
    set_one_h(c, run, one_h);
    HIPTRY(c, hipStreamSynchronize(c->stream));
    const int64_t n = s->n, o = R.p0;
    StateBuf& S = c->S[c->cur];
    int rc = 0;
    rc |= copy_col<uint64_t>(c, S.id + o, s->id, n, 0);
    rc |= copy_col<uint8_t>(c, S.mask + o, s->mask, n, 1);
    rc |= copy_col<uint8_t>(c, S.species + o, s->species, n, 0);
    std::vector<float> tmp;
    for (int a = 0; a < 3; ++a) {
        if (c->mixed) {
            rc |= copy_col<double>(c, S.pos64[a] + o, s->pos[a], n, 0.0);
        } else {
            tmp.assign(n, 0.0f);
            for (int64_t i = 0; i < n; ++i) tmp[i] = (float)s->pos[a][i];
            rc |= copy_col<float>(c, S.pos32[a] + o, tmp.data(), n, 0.0f);
            HIPTRY(c, hipStreamSynchronize(c->stream));
        }
        rc |= copy_col<float>(c, S.vel[a] + o, s->vel[a], n, 0.0f);
        if (c->has_mhd) rc |= copy_col<float>(c, S.B[a] + o, s->bfield[a], n, 0.0f);
    }
    rc |= copy_col<float>(c, S.mass + o, s->mass, n, 0.0f);
    rc |= copy_col<float>(c, S.h + o, s->smoothing_length, n, 0.0f);
    rc |= copy_col<float>(c, S.u + o, s->internal_energy, n, 0.0f);
    rc |= copy_col<float>(c, S.rho + o, s->density, n, 1.0f);
    if (c->has_mhd) rc |= copy_col<float>(c, S.psi + o, s->psi, n, 0.0f);
    if (rc) return SPH_E_DEVICE;
    // new initial data: the run restarts at t = 0 and takes its first dt from the next prime
    HIPTRY(c, launch_restart_run(c->runs_d, c->W, run, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));  // the columns above were copied from pageable memory
    if (c->binned) HIPTRY(c, launch_fill_u32(c->W.count, 0u, c->ncell_tot, c->stream));
    c->binned = false;
    c->sorted = false;
    c->dens_valid = false;
    return join_copies(c);  // earlier asynchronous copy-outs have landed too
}

int64_t sph_lattice_sites(const sph_lattice_block* blocks, int32_t n_blocks)
{
    return lattice_sites(blocks, n_blocks);
}

int sph_init_run(sph_ctx* c, int32_t run, const sph_lattice_block* blocks, int32_t n_blocks, uint64_t seed,
                 double jitter)
{
    if (!c || run < 0 || run >= c->nruns) return SPH_E_INVALID;
    const int64_t sites = lattice_sites(blocks, n_blocks);
    if (sites < 0) return fail(c, SPH_E_INVALID, "init_run: invalid lattice block (at most 8 blocks)");
    const RunDev& R = c->runs_h[run];
    if (sites != R.np) return fail(c, SPH_E_INVALID, "init_run: the blocks' lattice sites must equal the run's size");
    if (!(jitter >= 0.0 && jitter < 0.5)) return fail(c, SPH_E_INVALID, "init_run: jitter must be in [0, 0.5)");
    for (int b = 0; b < n_blocks; ++b)
        for (int a = 0; a < 3; ++a) {
            const double lo = R.lo[a], hi = R.lo[a] + R.L[a];
            if (blocks[b].lo[a] < lo || blocks[b].hi[a] > hi)
                return fail(c, SPH_E_INVALID, "init_run: a block extends outside the run's box");
        }
    const sph_run_config& cfg = c->cfg[run];
    HIPTRY(c, launch_init_lattice(blocks, n_blocks, R.np, R.p0, seed, jitter, cfg.h0, cfg.gamma, c->mixed ? 1 : 0,
                                  c->has_mhd ? 1 : 0, c->S[c->cur], c->stream));
    // as an upload: time restarts at 0 and the next step primes the run (enqueued, nothing waits)
    set_one_h(c, run, true);  // the fill gives every particle the run's h0
    HIPTRY(c, launch_restart_run(c->runs_d, c->W, run, c->stream));
    if (c->binned) HIPTRY(c, launch_fill_u32(c->W.count, 0u, c->ncell_tot, c->stream));
    c->binned = false;
    c->sorted = false;
    c->dens_valid = false;
    return SPH_OK;
}

int sph_host_alloc(int64_t bytes, void** out)
{
    if (!out || bytes < 1) return SPH_E_INVALID;
    *out = nullptr;
    return hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault) == hipSuccess ? SPH_OK : SPH_E_NOMEM;
}

int sph_host_free(void* p)
{
    if (!p) return SPH_OK;
    return hipHostFree(p) == hipSuccess ? SPH_OK : SPH_E_DEVICE;
}

int sph_download_async(sph_ctx* c, int32_t run, const sph_host_soa* d, sph_conserved* cons)
{
    if (!c || !d || run < 0 || run >= c->nruns) return SPH_E_INVALID;
    const RunDev& R = c->runs_h[run];
    if (d->n != R.np) return fail(c, SPH_E_INVALID, "download_async: n must equal the run's size");
    if (!c->mixed && (d->pos[0] || d->pos[1] || d->pos[2]))
        return fail(c, SPH_E_INVALID, "download_async: positions need conversion in FP32 contexts; use sph_download");
    const int64_t n = R.np, o = R.p0;
    const StateBuf& S = c->S[c->cur];
    struct Col {
        void* dst;
        const void* src;
        int64_t bytes;
    };
    std::vector<Col> cols;
    auto col = [&](void* dst, const void* src, size_t elem) {
        if (dst && src) cols.push_back({dst, src, (int64_t)elem * n});
    };
    col(d->id, S.id + o, 8);
    col(d->mask, S.mask + o, 1);
    col(d->species, S.species + o, 1);
    for (int a = 0; a < 3; ++a) {
        if (c->mixed) col(d->pos[a], S.pos64[a] + o, 8);
        col(d->vel[a], S.vel[a] + o, 4);
        col(d->dvdt[a], c->W.dvdt[a] + o, 4);
        if (c->has_mhd) {
            col(d->bfield[a], S.B[a] + o, 4);
            col(d->dbdt[a], c->W.dbdt[a] + o, 4);
        }
    }
    col(d->mass, S.mass + o, 4);
    col(d->density, S.rho + o, 4);
    col(d->smoothing_length, S.h + o, 4);
    col(d->internal_energy, S.u + o, 4);
    col(d->dudt, c->W.dudt + o, 4);
    col(d->vsig, c->W.vsig + o, 4);
    col(d->n_neighbours, c->W.nnb + o, 4);
    if (c->has_mhd) {
        col(d->psi, S.psi + o, 4);
        col(d->dpsidt, c->W.dpsidt + o, 4);
        col(d->divb, c->W.divb + o, 4);
    }
    // the FP64 sums and CFL dt that the run's last step (or reduction) left on the device
    static_assert(sizeof(sph_conserved) == 8 * sizeof(double), "sph_conserved is 8 doubles");
    if (cons) cols.push_back({cons, c->W.cons + (int64_t)run * 8, (int64_t)sizeof(sph_conserved)});

    // Pools snapshot the columns into their device ring on the context's stream (device-to-device,
    // at HBM speed) and copy the snapshot to the host on a second stream, so the copies over PCIe
    // overlap the steps that follow; otherwise, or when the ring is too small for the run, or while
    // the stream is being captured into a graph, the copies go out on the context's stream.
    int64_t need = 0;
    for (const Col& k : cols) need += (k.bytes + 255) & ~int64_t(255);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPTRY(c, hipStreamIsCapturing(c->stream, &cap));
    int64_t off = 0;
    int rc = SPH_OK;
    if (c->snap && cap == hipStreamCaptureStatusNone && need > 0 && snap_alloc(c, need, off, rc)) {
        if (!c->copy_stream) {
            HIPTRY(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
            HIPTRY(c, hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming));
        }
        int64_t at = off;
        for (const Col& k : cols) {
            HIPTRY(c, hipMemcpyAsync(c->snap + at, k.src, (size_t)k.bytes, hipMemcpyDeviceToDevice, c->stream));
            at += (k.bytes + 255) & ~int64_t(255);
        }
        HIPTRY(c, hipEventRecord(c->snap_ready, c->stream));
        HIPTRY(c, hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
        at = off;
        for (const Col& k : cols) {
            HIPTRY(c, hipMemcpyAsync(k.dst, c->snap + at, (size_t)k.bytes, hipMemcpyDeviceToHost, c->copy_stream));
            at += (k.bytes + 255) & ~int64_t(255);
        }
        hipEvent_t done = nullptr;
        if (!c->snap_free.empty()) {
            done = c->snap_free.back();
            c->snap_free.pop_back();
        } else {
            HIPTRY(c, hipEventCreateWithFlags(&done, hipEventDisableTiming));
        }
        HIPTRY(c, hipEventRecord(done, c->copy_stream));
        c->snap_q.push_back({off, need, done});
        return SPH_OK;
    }
    if (rc) return rc;
    for (const Col& k : cols)
        HIPTRY(c, hipMemcpyAsync(k.dst, k.src, (size_t)k.bytes, hipMemcpyDeviceToHost, c->stream));
    return SPH_OK;
}

int sph_download(sph_ctx* c, int32_t run, sph_host_soa* d)
{
    if (!c || !d || run < 0 || run >= c->nruns) return SPH_E_INVALID;
    const RunDev& R = c->runs_h[run];
    if (d->n != R.np) return fail(c, SPH_E_INVALID, "download: n must equal the run's capacity");
    const int64_t n = R.np, o = R.p0;
    const StateBuf& S = c->S[c->cur];
    int rc = 0;
    rc |= fetch_col(c, d->id, S.id + o, n);
    rc |= fetch_col(c, d->mask, S.mask + o, n);
    rc |= fetch_col(c, d->species, S.species + o, n);
    std::vector<float> tmp[3];
    for (int a = 0; a < 3; ++a) {
        if (c->mixed) rc |= fetch_col(c, d->pos[a], S.pos64[a] + o, n);
        else if (d->pos[a]) { tmp[a].resize(n); rc |= fetch_col(c, tmp[a].data(), S.pos32[a] + o, n); }
        rc |= fetch_col(c, d->vel[a], S.vel[a] + o, n);
        rc |= fetch_col(c, d->dvdt[a], c->W.dvdt[a] + o, n);
        if (c->has_mhd) {
            rc |= fetch_col(c, d->bfield[a], S.B[a] + o, n);
            rc |= fetch_col(c, d->dbdt[a], c->W.dbdt[a] + o, n);
        }
    }
    rc |= fetch_col(c, d->mass, S.mass + o, n);
    rc |= fetch_col(c, d->density, S.rho + o, n);
    rc |= fetch_col(c, d->smoothing_length, S.h + o, n);
    rc |= fetch_col(c, d->internal_energy, S.u + o, n);
    rc |= fetch_col(c, d->dudt, c->W.dudt + o, n);
    rc |= fetch_col(c, d->vsig, c->W.vsig + o, n);
    rc |= fetch_col(c, d->n_neighbours, c->W.nnb + o, n);
    if (c->has_mhd) {
        rc |= fetch_col(c, d->psi, S.psi + o, n);
        rc |= fetch_col(c, d->dpsidt, c->W.dpsidt + o, n);
        rc |= fetch_col(c, d->divb, c->W.divb + o, n);
    }
    if (rc) {
        hipStreamSynchronize(c->stream);  // copies already enqueued into the caller's arrays finish first
        return SPH_E_DEVICE;
    }
    HIPTRY(c, hipStreamSynchronize(c->stream));
    if (!c->mixed)
        for (int a = 0; a < 3; ++a)
            if (d->pos[a])
                for (int64_t i = 0; i < n; ++i) d->pos[a][i] = (double)tmp[a][i];
    return join_copies(c);  // earlier asynchronous copy-outs have landed too
}

int sph_stage_drift(sph_ctx* c, double dt)
{
    if (!c) return SPH_E_INVALID;
    return do_bin(c, dt, 1);
}

int sph_stage_resort(sph_ctx* c)
{
    if (!c) return SPH_E_INVALID;
    return do_resort(c);
}

int sph_stage_density(sph_ctx* c)
{
    if (!c) return SPH_E_INVALID;
    return do_density(c);
}

int sph_stage_neighbour_counts(sph_ctx* c)
{
    if (!c) return SPH_E_INVALID;
    return do_density(c);
}

int sph_stage_forces(sph_ctx* c)
{
    if (!c) return SPH_E_INVALID;
    return do_forces(c);
}

int sph_stage_integrate(sph_ctx* c, double dt)
{
    if (!c) return SPH_E_INVALID;
    return do_kick(c, dt, 1);
}

static int read_cons(sph_ctx* c, std::vector<double>& h)
{
    h.resize((size_t)c->nruns * 8);
    if (c->nruns == 0) return SPH_OK;
    HIPTRY(c, hipMemcpyAsync(h.data(), c->W.cons, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return SPH_OK;
}

int sph_reduce_min_dt(sph_ctx* c, double* dt)
{
    if (!c || !dt) return SPH_E_INVALID;
    int rc = do_kick(c, 0.0, 0);
    std::vector<double> h;
    if (!rc) rc = read_cons(c, h);
    if (rc) return rc;
    for (int r = 0; r < c->nruns; ++r) dt[r] = h[(size_t)r * 8 + 5];
    return SPH_OK;
}

int sph_reduce_conserved(sph_ctx* c, sph_conserved* o)
{
    if (!c || !o) return SPH_E_INVALID;
    int rc = do_kick(c, 0.0, 0);
    std::vector<double> h;
    if (!rc) rc = read_cons(c, h);
    if (rc) return rc;
    for (int r = 0; r < c->nruns; ++r) {
        const double* p = &h[(size_t)r * 8];
        o[r].mass = p[0];
        o[r].momentum[0] = p[1];
        o[r].momentum[1] = p[2];
        o[r].momentum[2] = p[3];
        o[r].energy = p[4];
        o[r].dt_next = p[5];
        o[r].time = p[6];
        o[r].c_h = p[7];
    }
    return SPH_OK;
}

int sph_prime(sph_ctx* c)
{
    if (!c) return SPH_E_INVALID;
    if (c->nruns == 0) return SPH_OK;
    int rc = do_resort(c);
    if (!rc) rc = do_density(c);
    if (!rc) rc = do_forces(c);
    if (!rc) rc = do_kick(c, 0.0, 0);  // sets the first dt of runs flagged by upload / set_runs
    if (!rc) HIPTRY(c, launch_fill_u32(reinterpret_cast<uint32_t*>(c->W.prime_flag), 0u, c->nruns, c->stream));
    return rc;
}

int sph_step(sph_ctx* c, int32_t n_steps)
{
    if (!c || n_steps < 0) return SPH_E_INVALID;
    if (c->nruns == 0) return SPH_OK;
    for (int s = 0; s < n_steps; ++s) {
        int rc = do_step(c);
        if (rc) return rc;
    }
    return SPH_OK;
}

int sph_graph_capture(sph_ctx* c)
{
    if (!c) return SPH_E_INVALID;
    invalidate_graph(c);
    // one step swaps the buffer sets, so a two-step graph returns to the starting parity
    const int parity = c->cur;
    const bool b0 = c->binned, s0 = c->sorted, d0 = c->dens_valid;
    c->binned = false;  // counts are zero at a step boundary
    HIPTRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = do_step(c);
    if (!rc) rc = do_step(c);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->cur = parity;
    c->binned = b0;
    c->sorted = s0;
    c->dens_valid = d0;
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(c, SPH_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    c->graph = g;
    HIPTRY(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
    c->graph_parity = parity;
    return SPH_OK;
}

int sph_graph_launch(sph_ctx* c, int32_t n_steps)
{
    if (!c || n_steps < 0) return SPH_E_INVALID;
    if (!c->gexec) return fail(c, SPH_E_STATE, "no captured graph");
    if (n_steps == 0) return SPH_OK;  // nothing ran: the stage state stays as it is
    if (c->binned) {
        HIPTRY(c, launch_fill_u32(c->W.count, 0u, c->ncell_tot, c->stream));
        c->binned = false;
    }
    int s = 0;
    if (c->cur != c->graph_parity && n_steps > 0) {
        int rc = do_step(c);
        if (rc) return rc;
        ++s;
    }
    for (; s + 2 <= n_steps; s += 2) HIPTRY(c, hipGraphLaunch(c->gexec, c->stream));
    c->sorted = true;
    c->dens_valid = true;
    c->binned = false;
    if (s < n_steps) return do_step(c);
    return SPH_OK;
}

int sph_get_sort(sph_ctx* c, int32_t run, uint32_t* key, uint32_t* cell_start)
{
    if (!c || run < 0 || run >= c->nruns) return SPH_E_INVALID;
    if (!c->sorted) return fail(c, SPH_E_STATE, "no resort since the last drift/upload");
    const RunDev& R = c->runs_h[run];
    std::vector<uint32_t> cs((size_t)R.ncell + 1);
    HIPTRY(c, hipMemcpyAsync(cs.data(), c->W.cs + R.c0, sizeof(uint32_t) * cs.size(), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    for (auto& v : cs) v -= (uint32_t)R.p0;
    if (cell_start) std::copy(cs.begin(), cs.end(), cell_start);
    if (key)
        for (int cc = 0; cc < R.ncell; ++cc)
            for (uint32_t i = cs[cc]; i < cs[cc + 1]; ++i) key[i] = (uint32_t)cc;
    return SPH_OK;
}

int sph_profile_begin(sph_ctx* c, int32_t max_steps)
{
    if (!c || max_steps < 1) return SPH_E_INVALID;
    for (auto e : c->prof_ev) hipEventDestroy(e);
    c->prof_ev.assign((size_t)max_steps * (SPH_NUM_STAGES + 1), nullptr);
    for (auto& e : c->prof_ev) HIPTRY(c, hipEventCreate(&e));
    c->prof_cap = max_steps;
    c->prof_steps = 0;
    c->profiling = true;
    return SPH_OK;
}

int sph_profile_end(sph_ctx* c, sph_profile* out, double* step_ms)
{
    if (!c || !out) return SPH_E_INVALID;
    if (!c->profiling) return fail(c, SPH_E_STATE, "profiling was not started");
    c->profiling = false;
    HIPTRY(c, hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof(*out));
    const int n = c->prof_steps;
    out->steps = n;
    std::vector<double> tot(n);
    for (int s = 0; s < n; ++s) {
        hipEvent_t* e = &c->prof_ev[(size_t)s * (SPH_NUM_STAGES + 1)];
        for (int k = 0; k < SPH_NUM_STAGES; ++k) {
            float ms = 0;
            HIPTRY(c, hipEventElapsedTime(&ms, e[k], e[k + 1]));
            out->stage_ms[k] += ms;
        }
        float ms = 0;
        HIPTRY(c, hipEventElapsedTime(&ms, e[0], e[SPH_NUM_STAGES]));
        tot[s] = ms;
        if (step_ms) step_ms[s] = ms;
    }
    if (n > 0) {
        for (int k = 0; k < SPH_NUM_STAGES; ++k) out->stage_ms[k] /= n;
        double sum = 0;
        for (double v : tot) sum += v;
        out->step_ms_mean = sum / n;
        std::vector<double> srt = tot;
        std::sort(srt.begin(), srt.end());
        auto pct = [&](double p) {
            double idx = p * (n - 1);
            int lo = (int)std::floor(idx), hi = std::min(n - 1, lo + 1);
            return srt[lo] + (srt[hi] - srt[lo]) * (idx - lo);
        };
        out->step_ms_p50 = pct(0.50);
        out->step_ms_p99 = pct(0.99);
    }
    return SPH_OK;
}

int sph_autotune(sph_ctx* c, int32_t n_steps, sph_tuning* best)
{
    if (!c || n_steps < 1) return SPH_E_INVALID;
    // device scratch of the search, released on every path
    struct Scratch {
        void* backup = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Scratch()
        {
            if (backup) hipFree(backup);
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
        }
    } sc;
    // snapshot the whole arena and the run configurations: timing leaves the state untouched
    HIPTRY(c, hipStreamSynchronize(c->stream));
    const int cur0 = c->cur;
    const size_t bytes = c->arena_bytes;
    HIPTRY(c, hipMalloc(&sc.backup, bytes));
    HIPTRY(c, hipEventCreate(&sc.e0));
    HIPTRY(c, hipEventCreate(&sc.e1));
    HIPTRY(c, hipMemcpyAsync(sc.backup, c->arena, bytes, hipMemcpyDeviceToDevice, c->stream));
    const std::vector<sph_run_config> cfg0 = c->cfg;
    const sph_tuning t0 = c->tune;

    // candidate launch shapes: chunked passes (pair_balance 0/1) and whole-cell passes (2), each at
    // supercell edges of 1 and 1.15 interaction radii, one wave per supercell; then the winner's
    // persistent-grid variants (waves per compute unit)
    std::vector<sph_tuning> cands;
    for (int dc : {128, 256})
        for (int fc : {64, 128})
            for (int k : {16, 64})
                for (int bal : {0, 1}) cands.push_back({dc, fc, 0, k, bal});
    for (int k : {16, 64}) cands.push_back({512, 128, 0, k, 2});
    // supercell edges of 1/2, 1 and 1.15 interaction radii, each no finer than the run's own setting;
    // a grid the context's supercell table cannot hold is skipped (a context reserves cells for its
    // runs' own grids and for edges of 1 H)
    std::vector<float> cfs;
    for (float cf : {0.5f, 1.0f, 1.15f}) {
        std::vector<sph_run_config> cfg = cfg0;
        int64_t cells = 0;
        for (auto& r : cfg) {
            r.cell_factor = (float)std::max<double>(cf, cell_factor_of(r.cell_factor));
            std::string why;
            RunDev R;
            cells += make_run(r, R, why) ? R.ncell : (1LL << 40);
        }
        bool dup = false;
        for (float prev : cfs) {
            bool same = true;
            for (const auto& r : cfg0)
                same = same && std::max<double>(prev, cell_factor_of(r.cell_factor)) ==
                                   std::max<double>(cf, cell_factor_of(r.cell_factor));
            dup = dup || same;
        }
        if (cells <= c->ncell_cap && !dup) cfs.push_back(cf);
    }
    const int wpcs[4] = {4, 8, 16, 32};

    double best_t = 1e300;
    sph_tuning best_tune = t0;
    float best_cf = -1.0f;
    auto time_one = [&](float cf, const sph_tuning& cand) -> int {
        for (int r = 0; r < c->nruns; ++r)
            c->cfg[r].cell_factor = (float)std::max<double>(cf, cell_factor_of(cfg0[r].cell_factor));
        if (!balance_allowed(c->cfg, cand.pair_balance)) return SPH_OK;  // not a candidate on this grid
        HIPTRY(c, hipMemcpyAsync(c->arena, sc.backup, bytes, hipMemcpyDeviceToDevice, c->stream));
        c->cur = cur0;
        c->tune = cand;
        int rc = regrid(c);
        if (rc) return rc;
        if ((rc = do_step(c))) return rc;  // warm-up
        HIPTRY(c, hipEventRecord(sc.e0, c->stream));
        for (int s = 0; s < n_steps; ++s)
            if ((rc = do_step(c))) return rc;
        HIPTRY(c, hipEventRecord(sc.e1, c->stream));
        HIPTRY(c, hipEventSynchronize(sc.e1));
        float ms = 0;
        HIPTRY(c, hipEventElapsedTime(&ms, sc.e0, sc.e1));
        if (ms < best_t) {
            best_t = ms;
            best_tune = cand;
            best_cf = cf;
        }
        return SPH_OK;
    };
    auto search = [&]() -> int {
        int rc = SPH_OK;
        for (float cf : cfs)
            for (const sph_tuning& cand : cands)
                if ((rc = time_one(cf, cand))) return rc;
        const sph_tuning shape = best_tune;
        const float shape_cf = best_cf;
        for (int w : wpcs) {
            sph_tuning cand = shape;
            cand.waves_per_cu = w;
            if ((rc = time_one(shape_cf, cand))) return rc;
        }
        return SPH_OK;
    };
    const int rc = search();

    // restore the state; apply the winner only if the search completed
    c->cfg = cfg0;
    if (!rc && best_cf > 0)
        for (int r = 0; r < c->nruns; ++r)
            c->cfg[r].cell_factor = (float)std::max<double>(best_cf, cell_factor_of(cfg0[r].cell_factor));
    c->tune = rc ? t0 : best_tune;
    c->cur = cur0;
    HIPTRY(c, hipMemcpyAsync(c->arena, sc.backup, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    const int rc2 = regrid(c);
    c->binned = false;  // counts were cleared by regrid
    c->sorted = false;
    c->dens_valid = false;
    if (best) *best = c->tune;
    return rc ? rc : rc2;
}

}  // extern "C"
