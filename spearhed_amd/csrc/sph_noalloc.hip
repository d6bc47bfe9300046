// sph_noalloc.hip -- sph_density_noalloc: the density pass (with its resort) for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). It runs the same kernels as sph_stage_resort + sph_stage_density on a context, on one
// run whose particles the caller holds in device columns, and writes density and neighbour counts
// in the caller's particle order.
//
// All working memory (run table, sorted copy of the particles, supercell histogram and starts,
// interaction records, permutation) lives in `scratch`, one file-scope device array per GPU that
// stays allocated until the process ends. Host code never passes the array's name to a kernel: the
// name is a host-side stand-in, not a device address. The device address is taken once per device
// with hipGetSymbolAddress and kept in `g_scratch`. Every call on a GPU shares its array, so calls
// that overlap on one GPU (different streams or threads) race; each call clears what must start at
// zero (the supercell histogram) with a fill kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "../../include/spearhed_sph.h"
#include "sph_device.hpp"
#include "sph_launch.hpp"

namespace sph {
bool make_run_config(const sph_run_config& c, RunDev& R, char* why, int why_len);
}

using namespace sph;

namespace {

constexpr int64_t kMaxParticles = SPH_NOALLOC_MAX_PARTICLES;
constexpr int64_t kMaxCells = SPH_NOALLOC_MAX_CELLS;

// Byte offsets of the scratch regions for kMaxParticles / kMaxCells (each 256-byte aligned).
struct Layout {
    size_t off = 0;
    constexpr size_t take(size_t bytes)
    {
        off = (off + 255) & ~size_t(255);
        const size_t at = off;
        off += bytes;
        return at;
    }
};

struct Regions {
    size_t run, blk, dt, time, chbits, prime, ids, zero, mask1, species0;
    size_t bid, bmask, bspecies, bpos[3], bvel[3], bmass, bh, bu, brho;
    size_t nnb, key, rank, tmp, tmpid, count, cs, bsum, p0, p1, p2, p3, perm;
    size_t total;
};

constexpr Regions regions()
{
    Regions g{};
    Layout l;
    const size_t N = (size_t)kMaxParticles, C = (size_t)kMaxCells;
    const size_t nkb = (N + 255) / 256;
    g.run = l.take(sizeof(RunDev));
    g.blk = l.take(4 * nkb);
    g.dt = l.take(8);
    g.time = l.take(8);
    g.chbits = l.take(4);
    g.prime = l.take(4);
    g.ids = l.take(8 * N);
    g.zero = l.take(4 * N);
    g.mask1 = l.take(N);
    g.species0 = l.take(N);
    g.bid = l.take(8 * N);
    g.bmask = l.take(N);
    g.bspecies = l.take(N);
    for (int a = 0; a < 3; ++a) g.bpos[a] = l.take(8 * N);  // FP64 (mixed) or FP32 (plain) positions
    for (int a = 0; a < 3; ++a) g.bvel[a] = l.take(4 * N);
    g.bmass = l.take(4 * N);
    g.bh = l.take(4 * N);
    g.bu = l.take(4 * N);
    g.brho = l.take(4 * N);
    g.nnb = l.take(4 * N);
    g.key = l.take(4 * N);
    g.rank = l.take(4 * N);
    g.tmp = l.take(4 * N);
    g.tmpid = l.take(8 * N);
    g.count = l.take(4 * C);
    g.cs = l.take(4 * (C + 1));
    g.bsum = l.take(4 * ((C + 2047) / 2048 + 1));
    g.p0 = l.take(16 * N);
    g.p1 = l.take(16 * N);
    g.p2 = l.take(16 * N);
    g.p3 = l.take(16 * N);
    g.perm = l.take(4 * N);
    g.total = (l.off + 255) & ~size_t(255);
    return g;
}

constexpr Regions kReg = regions();
static_assert(kReg.total <= SPH_NOALLOC_SCRATCH_BYTES, "the scratch layout exceeds its stated size");

}  // namespace

// The scratch array of every GPU this library runs on (its own copy per device).
static __device__ __align__(256) unsigned char scratch[kReg.total];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

__global__ void k_set_run(RunDev* dst, RunDev val) { *dst = val; }

__global__ __launch_bounds__(256) void k_iota64(uint64_t* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = (uint64_t)i;
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t* p, uint8_t v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// Sorted slot f came from input index perm[f]: results back into the caller's order.
__global__ __launch_bounds__(256) void k_unsort(const int32_t* __restrict__ perm, const float* __restrict__ rho,
                                                const uint32_t* __restrict__ nnb, const uint8_t* __restrict__ mask,
                                                const uint8_t* __restrict__ species, float* __restrict__ rho_out,
                                                uint32_t* __restrict__ nnb_out, int64_t n)
{
    const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (f >= n) return;
    const int32_t i = perm[f];
    const bool tgt = is_target(species[f], mask[f]);
    rho_out[i] = tgt ? rho[f] : 0.0f;
    nnb_out[i] = tgt ? nnb[f] : 0u;
}

int set_err(char* buf, int32_t len, int code, const char* msg)
{
    if (buf && len > 0) snprintf(buf, (size_t)len, "%s", msg);
    return code;
}

// Device address of this device's scratch array, looked up once per device (device current).
hipError_t scratch_ptr(unsigned char** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

inline unsigned grid256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int sph_density_noalloc(const sph_run_config* cfg, int32_t precision, const sph_device_soa* in,
                                   float* density_out, uint32_t* n_neighbours_out, void* stream, char* errbuf,
                                   int32_t errbuf_len)
{
    if (errbuf && errbuf_len > 0) errbuf[0] = '\0';
    // ---- host-side checks: nothing runs on the device unless every one passes
    if (!cfg || !in || !density_out || !n_neighbours_out)
        return set_err(errbuf, errbuf_len, SPH_E_INVALID, "density_noalloc: null configuration, input or output");
    if (precision != SPH_PREC_FP32 && precision != SPH_PREC_MIXED)
        return set_err(errbuf, errbuf_len, SPH_E_INVALID, "density_noalloc: unknown precision");
    const bool mixed = precision == SPH_PREC_MIXED;
    const int64_t n = in->n;
    if (n < 1) return set_err(errbuf, errbuf_len, SPH_E_INVALID, "density_noalloc: n must be >= 1");
    if (n > kMaxParticles) {
        char msg[160];
        snprintf(msg, sizeof msg, "density_noalloc: %lld particles exceed the maximum of %lld",
                 (long long)n, (long long)kMaxParticles);
        return set_err(errbuf, errbuf_len, SPH_E_CAPACITY, msg);
    }
    for (int a = 0; a < 3; ++a)
        if (mixed ? !in->pos[a] : !in->pos32[a])
            return set_err(errbuf, errbuf_len, SPH_E_INVALID,
                           mixed ? "density_noalloc: mixed precision needs the FP64 positions pos[3]"
                                 : "density_noalloc: FP32 precision needs the FP32 positions pos32[3]");
    if (!in->mass || !in->smoothing_length)
        return set_err(errbuf, errbuf_len, SPH_E_INVALID, "density_noalloc: mass and smoothing_length are required");
    RunDev R;
    char why[128];
    if (!make_run_config(*cfg, R, why, (int)sizeof why)) {
        char msg[192];
        snprintf(msg, sizeof msg, "density_noalloc: %s", why);
        return set_err(errbuf, errbuf_len, SPH_E_INVALID, msg);
    }
    if (R.ncell > kMaxCells) {
        char msg[160];
        snprintf(msg, sizeof msg, "density_noalloc: %d supercells exceed the maximum of %lld", R.ncell,
                 (long long)kMaxCells);
        return set_err(errbuf, errbuf_len, SPH_E_CAPACITY, msg);
    }
    R.p0 = 0;
    R.np = n;
    R.c0 = 0;
    R.kb_first = 0;
    R.kb_count = (int32_t)((n + 255) / 256);

    // ---- device
    unsigned char* base = nullptr;
    hipError_t e = scratch_ptr(&base);
    if (e != hipSuccess) return set_err(errbuf, errbuf_len, SPH_E_DEVICE, hipGetErrorString(e));
    hipStream_t s = (hipStream_t)stream;
    auto at = [&](size_t off) { return static_cast<void*>(base + off); };

    StateBuf A = {};  // the caller's columns (read only: the resort gathers them into B)
    A.id = static_cast<uint64_t*>(at(kReg.ids));  // ids 0..n-1: ties within a supercell in input order
    A.mask = in->mask ? const_cast<uint8_t*>(in->mask) : static_cast<uint8_t*>(at(kReg.mask1));
    A.species = in->species ? const_cast<uint8_t*>(in->species) : static_cast<uint8_t*>(at(kReg.species0));
    float* zero = static_cast<float*>(at(kReg.zero));
    for (int a = 0; a < 3; ++a) {
        if (mixed) A.pos64[a] = const_cast<double*>(in->pos[a]);
        else A.pos32[a] = const_cast<float*>(in->pos32[a]);
        A.vel[a] = zero;
    }
    A.mass = const_cast<float*>(in->mass);
    A.h = const_cast<float*>(in->smoothing_length);
    A.u = zero;
    A.rho = zero;

    StateBuf B = {};  // the sorted copy
    B.id = static_cast<uint64_t*>(at(kReg.bid));
    B.mask = static_cast<uint8_t*>(at(kReg.bmask));
    B.species = static_cast<uint8_t*>(at(kReg.bspecies));
    for (int a = 0; a < 3; ++a) {
        if (mixed) B.pos64[a] = static_cast<double*>(at(kReg.bpos[a]));
        else B.pos32[a] = static_cast<float*>(at(kReg.bpos[a]));
        B.vel[a] = static_cast<float*>(at(kReg.bvel[a]));
    }
    B.mass = static_cast<float*>(at(kReg.bmass));
    B.h = static_cast<float*>(at(kReg.bh));
    B.u = static_cast<float*>(at(kReg.bu));
    B.rho = static_cast<float*>(at(kReg.brho));

    WorkBuf W = {};
    W.nnb = static_cast<uint32_t*>(at(kReg.nnb));
    W.key = static_cast<uint32_t*>(at(kReg.key));
    W.rank = static_cast<uint32_t*>(at(kReg.rank));
    W.tmp = static_cast<uint32_t*>(at(kReg.tmp));
    W.tmpid = static_cast<uint64_t*>(at(kReg.tmpid));
    W.count = static_cast<uint32_t*>(at(kReg.count));
    W.cs = static_cast<uint32_t*>(at(kReg.cs));
    W.bsum = static_cast<uint32_t*>(at(kReg.bsum));
    W.P0 = static_cast<float4*>(at(kReg.p0));
    W.P1 = static_cast<float4*>(at(kReg.p1));
    W.P2 = static_cast<float4*>(at(kReg.p2));
    W.P3 = static_cast<float4*>(at(kReg.p3));
    W.dt = static_cast<double*>(at(kReg.dt));
    W.time = static_cast<double*>(at(kReg.time));
    W.chbits = static_cast<uint32_t*>(at(kReg.chbits));
    W.blk_run = static_cast<int32_t*>(at(kReg.blk));
    W.prime_flag = static_cast<int32_t*>(at(kReg.prime));
    W.perm = static_cast<int32_t*>(at(kReg.perm));
    // nbmask stays null: the density pass keeps no candidate masks (there is no force pass here)

    LaunchCtx L{};
    L.runs = static_cast<const RunDev*>(at(kReg.run));
    L.nruns = 1;
    L.ncell_tot = R.ncell;
    L.nkb = R.kb_count;
    L.mixed = mixed ? 1 : 0;
    L.has_mhd = 0;
    L.density_chunk = 128;  // a context's default tuning
    L.force_chunk = 64;
    L.kmax = 64;
    L.balance = 0;
    L.kern = cfg->kernel;
    L.rmax = R.reach;  // 2 for supercells finer than H: the 125-supercell stencil
    L.one_h = 0;
    L.persistent_waves = 0;
    L.S = A;
    L.W = W;
    L.stream = s;

    // per-call set-up, all by kernels on the caller's stream (nothing reads host memory later)
    k_set_run<<<1, 1, 0, s>>>(static_cast<RunDev*>(at(kReg.run)), R);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_fill_u32(W.count, 0u, R.ncell, s);  // the histogram starts at zero
    if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(W.blk_run), 0u, R.kb_count, s);
    if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(W.dt), 0u, 2, s);
    if (e == hipSuccess) e = launch_fill_u32(W.chbits, 0u, 1, s);
    if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(W.prime_flag), 0u, 1, s);
    if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(zero), 0u, n, s);
    if (e == hipSuccess) {
        k_iota64<<<grid256(n), 256, 0, s>>>(A.id, n);
        if (!in->mask) k_fill_u8<<<grid256(n), 256, 0, s>>>(A.mask, (uint8_t)1, n);
        if (!in->species) k_fill_u8<<<grid256(n), 256, 0, s>>>(A.species, (uint8_t)SPH_SPECIES_DEFAULT, n);
        e = hipGetLastError();
    }
    // the stages of sph_stage_resort + sph_stage_density
    if (e == hipSuccess) e = launch_drift_bin(L, 0.0, 0);
    if (e == hipSuccess) e = launch_scan(L);
    if (e == hipSuccess) e = launch_scatter_gather(L, B);
    if (e == hipSuccess) {
        L.S = B;
        e = launch_density(L);
    }
    if (e == hipSuccess) {
        k_unsort<<<grid256(n), 256, 0, s>>>(W.perm, B.rho, W.nnb, B.mask, B.species, density_out, n_neighbours_out,
                                            n);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return set_err(errbuf, errbuf_len, SPH_E_DEVICE, hipGetErrorString(e));
    return SPH_OK;
}
